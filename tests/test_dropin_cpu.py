"""The C++ drop-in layer without a GPU: the .prm parser against the
reference's own Src/FileParams.cpp (built into oracle/_ref by oracle/ref.mk
when the reference sources are present), the CLI's planner against the
oracle, and the reference's exit(1) paths."""
import json
import os
import subprocess

import pytest

from listctrecons_amd import scan
from oracle import oracle as O
from helpers import oracle_plan

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "listctrecons_amd", "bin", "LISTCTRecons")
REF_PARSER = os.path.join(ROOT, "oracle", "_ref", "ref_fileparams")


def run(*args):
    return subprocess.run(list(args), stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("scan"))
    w = scan.workload("tiny", view_offset=3, data_offset=5)
    return w, scan.write_scan_files(w, d)


def prm_variants(path, tmp):
    base = open(path).read()
    out = {"plain": base}
    out["tablefeed"] = base.replace("PitchValue:", "TableFeed:")
    out["extra_tokens"] = "Comment: hello\n" + base + "SmoothWindow: hann\nReconType: x\n"
    out["one_line"] = " ".join(base.split())
    paths = {}
    for k, v in out.items():
        p = os.path.join(tmp, k + ".prm")
        with open(p, "w") as f:
            f.write(v)
        paths[k] = p
    return paths


@pytest.mark.skipif(not os.path.exists(REF_PARSER), reason="reference parser not built")
def test_prm_parser_matches_reference(files, tmp_path):
    w, prm = files
    for name, p in prm_variants(prm, str(tmp_path)).items():
        ref = run(REF_PARSER, p)
        ours = run(CLI, "--prm", p)
        assert ref.returncode == 0 and ours.returncode == 0, name
        assert ours.stdout == ref.stdout, name


GOLDEN = json.load(open(os.path.join(ROOT, "tests", "golden", "prm_parser.json")))


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_prm_parser_matches_reference_vectors(name, tmp_path):
    """The reference's own FileParams.cpp output for 12 .prm files
    (tests/golden/make_prm_golden.py), replayed without the reference."""
    v = GOLDEN[name]
    p = tmp_path / (name + ".prm")
    with open(p, "w", newline="") as f:
        f.write(v["prm"])
    ours = run(CLI, "--prm", str(p))
    assert ours.returncode == v["returncode"]
    assert ours.stdout == v["stdout"]


@pytest.mark.skipif(not os.path.exists(REF_PARSER), reason="reference parser not built")
def test_missing_key_exits_like_reference(files, tmp_path):
    w, prm = files
    text = "".join(l for l in open(prm) if not l.startswith("Dz:"))
    p = str(tmp_path / "missing.prm")
    open(p, "w").write(text)
    ref = run(REF_PARSER, p)
    ours = run(CLI, "--prm", p)
    assert ref.returncode == ours.returncode == 1
    assert "Parameter not found in prm file:Dz" in ref.stdout
    assert ours.stdout == ref.stdout


def test_cli_plan_matches_oracle(files):
    w, prm = files
    out = run(CLI, "--plan", prm)
    assert out.returncode == 0, out.stdout
    pl, a, c = oracle_plan(w)
    line = [l for l in out.stdout.splitlines() if l.startswith("add ")][0].split()
    vals = dict(zip(line[0::2], line[1::2]))
    assert int(vals["add"]) == pl.ctg.add_projections
    assert int(vals["n_angle_add"]) == pl.ctg.n_angle_add
    assert int(vals["n_angle"]) == pl.ctg.n_angle
    assert int(vals["n_slices_recon"]) == pl.n_slices_recon
    blocks = [l.split() for l in out.stdout.splitlines() if l.startswith("block ")]
    for b in blocks:
        s, e, _, _ = O.block_info(pl, c, int(b[1]))
        assert (int(b[3]), int(b[5])) == (s, e)


def test_cli_errors(tmp_path):
    assert run(CLI).returncode == 1
    r = run(CLI, "--plan", str(tmp_path / "nope.prm"))
    assert r.returncode == 1 and "Parameter file not found" in r.stdout
    w = scan.workload("tiny", readings=300)
    prm = scan.write_scan_files(w, str(tmp_path / "short"))
    r = run(CLI, "--plan", prm)
    assert r.returncode == 1 and "outside of allowed data range" in r.stdout


@pytest.mark.parametrize("recon_type,keep", [
    ("Helicalfdkramprebingpu", 1000),                   # block 0's window is cut
    ("Helicalfdkramprebingpu-resident", None),          # half the file: later blocks cut
    ("Helicalfdkramprebingpu-resident-slab1-ckpt", None),
])
def test_cli_truncated_projection_file(tmp_path, recon_type, keep):
    """A projection file shorter than the planned windows exits 1 with a
    message before any device work (the reference read it unchecked,
    Src/DataReader.cpp:12-18)."""
    w = scan.workload("tiny", readings=1400, pitch=1.7)
    prm = scan.write_scan_files(w, str(tmp_path))
    raw = tmp_path / "proj.raw"
    size = raw.stat().st_size
    os.truncate(raw, keep if keep is not None else size // 2)
    r = run(CLI, prm, recon_type)
    assert r.returncode == 1, r.stdout[-1500:]
    assert "does not hold views" in r.stdout or "holds" in r.stdout, r.stdout[-1500:]
    assert "hip" not in r.stdout.lower()


def test_cli_refuses_a_foreign_checkpoint(tmp_path):
    """A streamed run with checkpointing checks "<output>.ckpt" before any
    device work and exits when it records another reconstruction."""
    w = scan.workload("tiny", readings=1400, pitch=1.7)
    prm = scan.write_scan_files(w, str(tmp_path))
    ck = tmp_path / "image.raw.ckpt"
    for text in ("LISTCTRecons checkpoint 1\nplan 9 9\ndone 1\n", "garbage\n"):
        ck.write_text(text)
        r = run(CLI, prm, "Helicalfdkramprebingpu-resident-ckpt")
        assert r.returncode == 1
        assert "belongs to another reconstruction" in r.stdout


REF_MAIN = os.path.join(ROOT, "oracle", "_ref", "ref_main")


@pytest.mark.skipif(not os.path.exists(REF_MAIN), reason="reference main.cpp not built")
def test_reference_main_binds_unchanged(files, tmp_path):
    """The reference's own Src/main.cpp, compiled unchanged against
    include/listctrecons and linked to libListCTRecons + liblct in place of its
    CUDA objects (oracle/ref.mk), exits like this package's CLI: usage without
    arguments, the parser's message and exit 1 on a missing key or file."""
    w, prm = files
    ref, ours = run(REF_MAIN), run(CLI)
    assert ref.returncode == ours.returncode == 1
    assert ref.stdout == ours.stdout == "Usage LISTCTRecons <PRM filename>\n"
    text = "".join(l for l in open(prm) if not l.startswith("NProjTurn:"))
    p = str(tmp_path / "missing.prm")
    open(p, "w").write(text)
    ref, ours = run(REF_MAIN, p), run(CLI, p)
    assert ref.returncode == ours.returncode == 1
    assert "Parameter not found in prm file:NProjTurn" in ref.stdout
    assert ref.stdout == ours.stdout
    ref, ours = run(REF_MAIN, str(tmp_path / "nope.prm")), run(CLI, str(tmp_path / "nope.prm"))
    assert ref.returncode == ours.returncode == 1
    assert ref.stdout == ours.stdout
    # a scan too short for the planner: the same message from ReconData
    ws = scan.workload("tiny", readings=300)
    short = scan.write_scan_files(ws, str(tmp_path / "short"))
    ref, ours = run(REF_MAIN, short), run(CLI, short)
    assert ref.returncode == ours.returncode == 1
    assert "outside of allowed data range" in ref.stdout
    assert ref.stdout == ours.stdout
