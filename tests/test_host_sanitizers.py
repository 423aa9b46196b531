"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer.

The C++ drop-in (.prm parser, planner, CLI, file and checkpoint checks) runs
over the CLI paths that need no GPU: the 12 golden .prm files, --plan,
missing files, a truncated projection file and foreign or damaged
checkpoints.  The FP64 oracle runs every one of its functions on a tiny scan.
Any sanitizer report fails the test."""
import json
import os
import shutil
import subprocess

import pytest

from listctrecons_amd import scan

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "listctrecons_amd")
LIB = os.path.join(PKG, "liblct.so")


@pytest.fixture(scope="module")
def cli(tmp_path_factory):
    if not os.path.exists(LIB):
        pytest.skip("liblct.so not built")
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("asan")
    exe = str(d / "cli_asan")
    cpp = os.path.join(PKG, "cpp")
    srcs = sorted(os.path.join(cpp, f) for f in os.listdir(cpp) if f.endswith(".cpp"))
    # the planner is host code of liblct: sanitize it too (it takes precedence
    # over the library's copy)
    srcs.append(os.path.join(PKG, "csrc", "lct_plan.cpp"))
    cmd = ["g++", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
           "-std=c++17", f"-I{os.path.join(ROOT, 'include')}",
           f"-I{os.path.join(ROOT, 'include', 'listctrecons')}", *srcs, "-o", exe,
           f"-L{PKG}", "-llct", f"-Wl,-rpath,{PKG}"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-3000:]
    return exe


def run(exe, *args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, *args], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, env=env)
    assert "Sanitizer" not in r.stdout and "runtime error" not in r.stdout, r.stdout[-3000:]
    return r


def test_parser_on_golden_prm_files(cli, tmp_path):
    golden = json.load(open(os.path.join(ROOT, "tests", "golden", "prm_parser.json")))
    for name, v in golden.items():
        p = tmp_path / (name + ".prm")
        with open(p, "w", newline="") as f:
            f.write(v["prm"])
        r = run(cli, "--prm", str(p))
        assert r.returncode == v["returncode"] and r.stdout == v["stdout"], name


def test_planner_and_error_exits(cli, tmp_path):
    w = scan.workload("tiny", view_offset=3, data_offset=5)
    prm = scan.write_scan_files(w, str(tmp_path / "scan"))
    assert run(cli, "--plan", prm).returncode == 0
    assert run(cli).returncode == 1
    assert run(cli, "--plan", str(tmp_path / "nope.prm")).returncode == 1


def test_truncated_file_and_checkpoints(cli, tmp_path):
    w = scan.workload("tiny", readings=1400, pitch=1.7)
    prm = scan.write_scan_files(w, str(tmp_path))
    raw = tmp_path / "proj.raw"
    os.truncate(raw, raw.stat().st_size // 2)
    for recon_type in ("Helicalfdkramprebingpu-resident",
                       "Helicalfdkramprebingpu-resident-slab1-ckpt"):
        assert run(cli, prm, recon_type).returncode == 1
    ck = tmp_path / "image.raw.ckpt"
    for text in ("LISTCTRecons checkpoint 1\nplan 9 9\ndone 1\n", "garbage\n", ""):
        ck.write_text(text)
        assert run(cli, prm, "Helicalfdkramprebingpu-resident-ckpt").returncode == 1


def test_oracle_under_sanitizers(tmp_path):
    """Every oracle function on a tiny scan (tests/native/oracle_sanitize.c),
    the oracle compiled from its source with ASan and UBSan (and OpenMP)."""
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    exe = str(tmp_path / "oracle_san")
    cmd = ["gcc", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
           "-fopenmp", "-std=c11", f"-I{os.path.join(ROOT, 'oracle')}",
           os.path.join(ROOT, "tests", "native", "oracle_sanitize.c"),
           os.path.join(ROOT, "oracle", "lct_oracle.c"), "-o", exe, "-lm"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-3000:]
    env = dict(os.environ, OMP_NUM_THREADS="4", ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       env=env, timeout=600)
    assert "Sanitizer" not in r.stdout and "runtime error" not in r.stdout, r.stdout[-3000:]
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout[-3000:]
