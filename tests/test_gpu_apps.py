"""The drop-in executables end to end on the GPU: each technique executable
reads a reference-format .bin, runs on the device and writes the result file,
which must be byte-identical to the golden file of the reference's writers
(single-run order, io.hpp:91-102; multi-run order for --partition 3 and 5,
io.hpp:104-136).  The golden files come from tests/golden/make_fixtures.py
and are pinned by the CPU executables in test_tools.py."""
import filecmp
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
BIN = os.path.join(ROOT, "set_intersection_amd", "bin")
FIXTURES = ["closed_form", "random_k12", "random_k23"]


def run_exe(exe, name, out, *extra):
    args = [os.path.join(BIN, exe), "--input", os.path.join(GOLDEN, name + ".bin"), "--output", str(out), *extra]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (args, r.stdout[-2000:], r.stderr[-2000:])
    return r


def golden(name, partition):
    return os.path.join(GOLDEN, name + ("" if partition is None else f"_p{partition}") + ".txt")


@pytest.mark.parametrize("exe,flags", [("ip", []), ("obs", []), ("obs", ["--warp", "--path"]), ("hi", []),
                                       ("hi", ["--warp", "--path"]), ("bitmap_dynamic", [])])
@pytest.mark.parametrize("partition", [None, 3])
@pytest.mark.parametrize("name", FIXTURES)
def test_technique_executables(exe, flags, partition, name, tmp_path):
    out = tmp_path / "r.txt"
    extra = list(flags) + ([] if partition is None else ["--partition", str(partition)])
    r = run_exe(exe, name, out, *extra)
    assert "Dataset characteristics" in r.stdout
    # the reference's device-timer rows for the upload (ip.cu:128-137)
    assert "Transfer to device" in r.stdout and "Compute set offsets" in r.stdout
    assert filecmp.cmp(str(out), golden(name, partition), shallow=False)


def test_upload_events(ctx):
    """si_dataset_upload_events times the copies and the offsets scan apart."""
    import set_intersection_amd as si
    from set_intersection_amd import datasets
    ds = ctx.upload(datasets.zipf_collection(3000, 1.2, 4, 3000, 1 << 16, 2))
    xfer, scan = ds.upload_ms()
    assert xfer > 0.0 and scan > 0.0
    ds.free()


@pytest.mark.parametrize("name", FIXTURES)
def test_bitmap_naive_executable(name, tmp_path):
    out = tmp_path / "r.txt"
    run_exe("bitmap_naive", name, out)
    assert filecmp.cmp(str(out), golden(name, None), shallow=False)


@pytest.mark.parametrize("variant", ["auto", "fp4", "i8", "f32", "popcount"])
@pytest.mark.parametrize("partition", [3, 5])
@pytest.mark.parametrize("name", ["random_k12", "random_k23"])
def test_mm_executable(variant, partition, name, tmp_path):
    """The dense-matrix technique on the fixtures with small universes (the
    reference's mm allocates partition x (universe + 1) floats)."""
    out = tmp_path / "r.txt"
    run_exe("mm", name, out, "--partition", str(partition), "--variant", variant)
    assert filecmp.cmp(str(out), golden(name, partition), shallow=False)


@pytest.mark.parametrize("variant", ["fp4", "f32"])
def test_mm_executable_refuses_fp32_range(variant, tmp_path):
    """Universe 2^32 - 1: the FP32-accumulated variants refuse (counts could
    round) and the executable reports the error."""
    args = [os.path.join(BIN, "mm"), "--input", os.path.join(GOLDEN, "closed_form.bin"), "--output",
            str(tmp_path / "r.txt"), "--partition", "3", "--variant", variant]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "2^24" in r.stdout + r.stderr


@pytest.mark.parametrize("partition", [None, 3])
@pytest.mark.parametrize("name", FIXTURES)
def test_misi_batched_executable(partition, name, tmp_path):
    out = tmp_path / "r.txt"
    extra = ["--no-graph"] if partition == 3 else []
    if partition is not None:
        extra += ["--partition", str(partition)]
    run_exe("misi_batched", name, out, *extra)
    assert filecmp.cmp(str(out), golden(name, partition), shallow=False)


@pytest.fixture(scope="module")
def check_writers(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("cw") / "check_writers")
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "set_intersection_amd", "host"), "-o",
                        exe, os.path.join(ROOT, "tests", "cpp", "check_writers.cpp")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    return exe


@pytest.mark.parametrize("exe", ["ip", "obs"])
def test_pipelined_many_runs(exe, check_writers, tmp_path):
    """28 runs (2000 Zipf sets, partition 300) through the executables' transfer
    and writing pipeline: the file equals the one the host writers produce from
    counts computed on the CPU (tests/cpp/check_writers.cpp)."""
    from set_intersection_amd import datasets
    coll = datasets.zipf_collection(2000, 1.2, 4, 3000, 1 << 16, 21)
    src = str(tmp_path / "z.bin")
    datasets.write_bin(coll, src)
    exp = str(tmp_path / "exp.txt")
    r = subprocess.run([check_writers, src, exp, "300"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = tmp_path / "got.txt"
    args = [os.path.join(BIN, exe), "--input", src, "--output", str(out), "--partition", "300"]
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    assert filecmp.cmp(str(out), exp, shallow=False)


def test_misi_batched_tune_file(check_writers, tmp_path):
    """misi_batched --tune-file: a tuned run writes the dataset's record
    (shape features + tuning), a second run uses the stored record, a
    collection of similar shape (other seed) starts from the nearest record,
    and a damaged store falls back to the defaults; every result file equals
    the CPU writer's."""
    from set_intersection_amd import datasets
    store = str(tmp_path / "tune.json")
    exe = os.path.join(BIN, "misi_batched")

    def case(seed, *flags):
        src = str(tmp_path / f"z{seed}.bin")
        if not os.path.exists(src):
            datasets.write_bin(datasets.zipf_collection(700, 1.2, 16, 4096, 1 << 16, seed), src)
        exp = str(tmp_path / f"exp{seed}.txt")
        if not os.path.exists(exp):
            r = subprocess.run([check_writers, src, exp, "0"], capture_output=True, text=True, timeout=300)
            assert r.returncode == 0, r.stderr
        out = tmp_path / "got.txt"
        r = subprocess.run([exe, "--input", src, "--output", str(out), "--partition", "100000", *flags],
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
        assert filecmp.cmp(str(out), exp, shallow=False)
        return r.stdout

    assert "autotuned" in case(1, "--tune", "--tune-file", store)
    text = open(store).read()
    assert '"features"' in text and '"tuning"' in text and '"deg2"' in text
    assert "stored" in case(1, "--tune-file", store) and "nearest" not in case(1, "--tune-file", store)
    assert "nearest stored" in case(2, "--tune-file", store)
    assert "autotuned from the stored record" in case(2, "--tune", "--tune-file", store)
    assert open(store).read().count('"k=700') == 2  # one record per shape
    with open(store, "w") as f:
        f.write('{"k=700;ubits=17": {"features": {"k": 7')  # truncated
    out = case(1, "--tune-file", store)
    assert "default" in out
