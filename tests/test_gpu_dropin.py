"""The LISTCTRecons CLI end to end on the GPU: reference file formats in,
image.raw and rebin.raw out, compared with the FP64 oracle."""
import os
import subprocess

import numpy as np
import pytest

from listctrecons_amd import scan
from oracle import oracle as O
from helpers import cpu_fan, hu_errors, oracle_plan, rel_rmse

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "listctrecons_amd", "bin", "LISTCTRecons")


@pytest.fixture(scope="module")
def case(tmp_path_factory):
    w = scan.workload("helical_16x368_256x32", view_offset=2, data_offset=3)
    fan = cpu_fan(w)
    d = str(tmp_path_factory.mktemp("cli"))
    prm = scan.write_scan_files(w, d, fan)
    pl, a, c = oracle_plan(w)
    ref = O.reconstruct(pl, fan, a, c, window=scan.WORKLOADS[w.name].window)
    return w, d, prm, fan, pl, a, c, ref


@pytest.mark.parametrize("recon_type,kind", [
    ("Helicalfdkramprebingpu-hann", "fp32"),
    ("Helicalfdkramprebingpu-hann-resident", "fp32"),
    ("Helicalfdkramprebingpu-hann-resident-x6", "fp32"),
    ("Helicalfdkramprebingpu-hann-resident-f16", "mixed"),
])
def test_cli_reconstructs_like_oracle(case, recon_type, kind):
    w, d, prm, fan, pl, a, c, ref = case
    print(f"CLI {recon_type}", flush=True)
    r = subprocess.run([CLI, prm, recon_type], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:]
    img = np.fromfile(os.path.join(d, "image.raw"), np.float32).reshape(ref.shape)
    if kind == "fp32":
        assert rel_rmse(img, ref) <= 1e-4
    else:
        assert hu_errors(img, ref)[0] <= 1.0


@pytest.mark.parametrize("recon_type,tol", [
    ("Helicalfdkramprebingpu-hann", 2e-6),
    ("Helicalfdkramprebingpu-hann-resident", 2e-6),       # FP32 direct filter into Q
    ("Helicalfdkramprebingpu-hann-resident-x6", 2e-6),    # BF16x6 MFMA filter, f32 storage
    ("Helicalfdkramprebingpu-hann-resident-f16", 2e-3),   # BF16x3 MFMA filter, f16 storage
])
def test_cli_rebin_raw_is_last_filtered_block(case, recon_type, tol):
    w, d, prm, fan, pl, a, c, ref = case
    rb = os.path.join(d, "rebin.raw")
    if os.path.exists(rb):
        os.remove(rb)
    print(f"rebin.raw {recon_type}", flush=True)
    r = subprocess.run([CLI, prm, recon_type], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:]
    g = pl.ctg
    got = np.fromfile(rb, np.float32).reshape(g.n_angle, g.nt, g.ns)
    s, _, _, _ = O.block_info(pl, c, pl.n_blocks - 1)
    first = s + pl.data_offset
    par = O.rebin(g, fan[first:first + g.n_angle_add], kind=1)
    q = O.filter_rows(O.ramp_taps(g.ns, g.ds, 1), par)
    assert rel_rmse(got, q) < tol


@pytest.mark.parametrize("recon_type,tol", [("Helicalfdkpioriramprebingpu", 1e-4),
                                             ("Helicalfdkpioriramprebingpu-resident", 1e-4),
                                             ("Helicalfdkpioriramprebingpu-f16", 1e-3)])
def test_cli_pi_original_stage_set(case, tmp_path, recon_type, tol):
    """The PI-original stage set from the factory string: the per-block loop,
    and the resident pipeline with the folded PI-original backprojector (FP32
    and f16 storage), against the oracle's PI-original reconstruction."""
    w, d, prm, fan, pl, a, c, ref = case
    print(f"CLI {recon_type}", flush=True)
    r = subprocess.run([CLI, prm, recon_type], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:]
    img = np.fromfile(os.path.join(d, "image.raw"), np.float32).reshape(ref.shape)
    ref_pi = O.reconstruct(pl, fan, a, c, rebin_kind=2, weight_kind=1, window=0, bp_kind=1)
    err = rel_rmse(img, ref_pi)
    print(f"  rel rmse {err:.2e}", flush=True)
    assert err <= tol


def test_factory_rejects_strings_starting_with_fdk(case):
    w, d, prm, fan, pl, a, c, ref = case
    r = subprocess.run([CLI, prm, "fdkHelical"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    assert r.returncode == 1
    assert "Error happens when building reconstructor." in r.stdout


@pytest.fixture(scope="module")
def three_blocks(tmp_path_factory):
    w = scan.workload("tiny", readings=1400, pitch=1.7, data_offset=4)
    fan = cpu_fan(w)
    dirs = {}
    for k in ("whole", "slab2", "resumed"):
        d = str(tmp_path_factory.mktemp(k))
        dirs[k] = (d, scan.write_scan_files(w, d, fan))
    pl, a, c = oracle_plan(w)
    assert pl.n_blocks == 3
    ref = O.reconstruct(pl, fan, a, c, window=w.window)
    return w, dirs, ref


def test_cli_slab_streaming_and_resume(three_blocks):
    """Streaming by slabs, and a checkpointed run stopped after one slab and
    resumed, write the same image.raw bit for bit as the one-shot resident
    run; it matches the oracle."""
    w, dirs, ref = three_blocks

    def cli(key, *extra):
        d, prm = dirs[key]
        print(f"CLI {key} {extra}", flush=True)
        return subprocess.run([CLI, prm, *extra], stdout=subprocess.PIPE,
                              stderr=subprocess.STDOUT, text=True, timeout=300)

    def image(key):
        return np.fromfile(os.path.join(dirs[key][0], "image.raw"), np.float32)

    r = cli("whole", "Helicalfdkramprebingpu-resident")
    assert r.returncode == 0, r.stdout[-2000:]
    whole = image("whole")
    assert rel_rmse(whole.reshape(ref.shape), ref) <= 1e-4

    r = cli("slab2", "Helicalfdkramprebingpu-resident-slab2")
    assert r.returncode == 0, r.stdout[-2000:]
    assert np.array_equal(image("slab2"), whole)

    ck = os.path.join(dirs["resumed"][0], "image.raw.ckpt")
    r = cli("resumed", "Helicalfdkramprebingpu-resident-ckpt", "--stop-after-slabs", "1")
    assert r.returncode == 3, r.stdout[-2000:]
    assert open(ck).read().splitlines()[-1] == "done 1"
    r = cli("resumed", "Helicalfdkramprebingpu-resident-ckpt")
    assert r.returncode == 0, r.stdout[-2000:]
    assert "Resuming after block 0" in r.stdout
    assert not os.path.exists(ck)
    assert np.array_equal(image("resumed"), whole)


def test_cli_checkpoint_of_another_scan_is_refused(three_blocks, tmp_path):
    """A checkpoint records the plan's geometry, the tube-angle and couch
    tables and the projection file (path and size): a stopped run's checkpoint
    is not resumed after the angle file changed (same sizes), and is resumed
    once it is restored."""
    w, dirs, ref = three_blocks
    fan = np.fromfile(os.path.join(dirs["resumed"][0], "proj.raw"), np.float32)
    d = str(tmp_path)
    prm = scan.write_scan_files(w, d, fan.reshape(-1, w.nt, w.ns))
    run = lambda *extra: subprocess.run([CLI, prm, "Helicalfdkramprebingpu-resident-ckpt", *extra],
                                        stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                        text=True, timeout=300)
    r = run("--stop-after-slabs", "1")
    assert r.returncode == 3, r.stdout[-2000:]
    ang_path = os.path.join(d, "angles.raw")
    ang = np.fromfile(ang_path, np.float32)
    (ang + np.float32(1e-3)).tofile(ang_path)
    r = run()
    assert r.returncode == 1, r.stdout[-2000:]
    assert "belongs to another reconstruction" in r.stdout
    ang.tofile(ang_path)
    r = run()
    assert r.returncode == 0, r.stdout[-2000:]
    assert "Resuming after block 0" in r.stdout
    img = np.fromfile(os.path.join(d, "image.raw"), np.float32).reshape(ref.shape)
    assert rel_rmse(img, ref) <= 1e-4


def test_gpu_memory_ctrl_semantics():
    """GPUMemoryCtrl<T> of the drop-in headers on the device (program built
    from tests/native/gpu_memory_ctrl.cpp by the package build)."""
    exe = os.path.join(ROOT, "tests", "native", "bin", "gpu_memory_ctrl")
    assert os.path.exists(exe), "build the package first"
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout[-2000:]


REF_MAIN = os.path.join(ROOT, "oracle", "_ref", "ref_main")


@pytest.mark.skipif(not os.path.exists(REF_MAIN), reason="reference main.cpp not built")
def test_reference_main_reconstructs_like_the_cli(case):
    """The reference's own Src/main.cpp, compiled unchanged against the drop-in
    headers and linked to this package (oracle/ref.mk), runs its hard-coded
    "Helicalfdkramprebingpu" reconstruction: image.raw and rebin.raw equal
    this package's CLI with the same string byte for byte, and the image is
    the FP64 oracle's Ram-Lak reconstruction to 1e-4."""
    w, d, prm, fan, pl, a, c, ref = case
    out = {}
    for name, cmd in (("reference main", [REF_MAIN, prm]),
                      ("cli", [CLI, prm, "Helicalfdkramprebingpu"])):
        for f in ("image.raw", "rebin.raw"):
            if os.path.exists(os.path.join(d, f)):
                os.remove(os.path.join(d, f))
        print(name, flush=True)
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stdout[-2000:]
        out[name] = tuple(open(os.path.join(d, f), "rb").read() for f in ("image.raw", "rebin.raw"))
    assert out["reference main"] == out["cli"]
    img = np.frombuffer(out["cli"][0], np.float32).reshape(ref.shape)
    ref_ramlak = O.reconstruct(pl, fan, a, c, window=0)
    assert rel_rmse(img, ref_ramlak) <= 1e-4
