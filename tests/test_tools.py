"""CPU tests of the host side: .bin reader/writers, result-file writers,
dataset tools (txt2bin, bin2txt, generate_dataset) and the CPU executables'
command-line contract.  No GPU."""
import filecmp
import os
import subprocess

import numpy as np
import pytest

from set_intersection_amd import datasets

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
BIN = os.path.join(ROOT, "set_intersection_amd", "bin")
OBIN = os.path.join(ROOT, "oracle", "bin")
HOST = os.path.join(ROOT, "set_intersection_amd", "host")
FIXTURES = ["closed_form", "random_k12", "random_k23"]


def run(args, **kw):
    return subprocess.run(args, capture_output=True, text=True, timeout=120, **kw)


@pytest.fixture(scope="module")
def check_writers(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("cw") / "check_writers")
    r = run(["g++", "-O2", "-std=c++17", "-I" + HOST, "-o", exe, os.path.join(ROOT, "tests", "cpp", "check_writers.cpp")])
    assert r.returncode == 0, r.stderr
    return exe


@pytest.mark.parametrize("name", FIXTURES)
def test_bin_roundtrip_python(name, tmp_path):
    c = datasets.read_bin(os.path.join(GOLDEN, name + ".bin"))
    p = str(tmp_path / "x.bin")
    datasets.write_bin(c, p)
    assert filecmp.cmp(p, os.path.join(GOLDEN, name + ".bin"), shallow=False)


def test_truncated_bin_rejected(tmp_path):
    data = open(os.path.join(GOLDEN, "random_k12.bin"), "rb").read()
    p = tmp_path / "t.bin"
    p.write_bytes(data[:-4])
    with pytest.raises(ValueError):
        datasets.read_bin(str(p))
    r = run([os.path.join(BIN, "bin2txt"), str(p), str(tmp_path / "o")])
    assert r.returncode == 1 and "truncated" in r.stdout


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("partition", [0, 3, 5])
def test_result_writers_match_golden(check_writers, name, partition, tmp_path):
    out = str(tmp_path / "r.txt")
    r = run([check_writers, os.path.join(GOLDEN, name + ".bin"), out, str(partition)])
    assert r.returncode == 0, r.stderr
    golden = os.path.join(GOLDEN, name + (".txt" if partition == 0 else f"_p{partition}.txt"))
    assert filecmp.cmp(out, golden, shallow=False)
    if partition:
        r = run([check_writers, os.path.join(GOLDEN, name + ".bin"), out, str(partition), "mm"])
        assert r.returncode == 0 and filecmp.cmp(out, golden, shallow=False)


@pytest.mark.parametrize("exe", ["stdintersection", "serial_intersect", "simd", "boost_bitset"])
@pytest.mark.parametrize("name", FIXTURES)
def test_cpu_executables_match_golden(exe, name, tmp_path):
    out = str(tmp_path / "r.txt")
    r = run([os.path.join(OBIN, exe), "--input", os.path.join(GOLDEN, name + ".bin"), "--output", out, "--threads", "3"])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Dataset characteristics" in r.stdout
    assert filecmp.cmp(out, os.path.join(GOLDEN, name + ".txt"), shallow=False)


def test_cli_contract():
    exe = os.path.join(OBIN, "stdintersection")
    r = run([exe])
    assert r.returncode == 1 and "No input dataset given! Exiting..." in r.stdout
    r = run([exe, "--bogus", "1"])
    assert r.returncode == 1 and "does not exist" in r.stdout
    r = run([exe, "--input"])
    assert r.returncode == 1 and "missing an argument" in r.stdout
    r = run([exe, "--help"])
    assert r.returncode == 0 and "--threads" in r.stdout


def test_txt2bin_bin2txt_roundtrip(tmp_path):
    txt = tmp_path / "in.txt"
    sets = [[1, 5, 9], [2, 3], [], [0, 2**31, 2**32 - 1], [7]]
    txt.write_text("\n".join(" ".join(map(str, s)) for s in sets) + "\n")
    r = run([os.path.join(BIN, "txt2bin"), str(txt), str(tmp_path / "d")])
    assert r.returncode == 0, r.stdout
    asc = datasets.read_bin(str(tmp_path / "d_asc.bin"))
    desc = datasets.read_bin(str(tmp_path / "d_desc.bin"))
    nonempty = [s for s in sets if s]  # empty lines are skipped
    assert asc.k == desc.k == len(nonempty)
    assert list(asc.sizes) == sorted(len(s) for s in nonempty)
    assert list(desc.sizes) == sorted((len(s) for s in nonempty), reverse=True)
    assert asc.universe == desc.universe == 2**32 - 1
    assert sorted(tuple(desc.set(i)) for i in range(desc.k)) == sorted(tuple(s) for s in nonempty)
    r = run([os.path.join(BIN, "bin2txt"), str(tmp_path / "d_desc.bin"), str(tmp_path / "back")])
    assert r.returncode == 0
    lines = (tmp_path / "back.txt").read_text().splitlines()
    assert [list(map(int, l.split())) for l in lines] == [list(desc.set(i)) for i in range(desc.k)]
    assert lines[0].endswith(" ")  # each value followed by a space, as bin2txt.cpp:33


@pytest.mark.parametrize("dist,extra", [("uniform", []), ("normal", ["--stdev", "50"]),
                                        ("bernoulli", ["--prob", "0.01"]), ("zipf", ["--skew", "0.9"])])
def test_generate_dataset_modes(dist, extra, tmp_path):
    pre = str(tmp_path / "g")
    args = [os.path.join(BIN, "generate_dataset"), "--distribution", dist, "--universe", "100000", "--k", "20",
            "--iterations", "3000", "--seed", "7", "--output", pre] + extra
    r = run(args)
    assert r.returncode == 0, r.stdout
    files = sorted(os.listdir(tmp_path))
    assert len(files) == 2 and files[0].startswith(f"g_asc_{dist}_20_100000") and files[1].startswith("g_desc_")
    c = datasets.read_bin(str(tmp_path / files[0]))
    assert c.k == 20 and c.universe == 100000
    for i in range(c.k):
        s = c.set(i).astype(np.int64)
        assert np.all(np.diff(s) > 0) and (s.size == 0 or (s[0] >= 0 and s[-1] <= 100000))
    assert list(c.sizes) == sorted(c.sizes)
    # reproducible with a seed
    args2 = list(args)
    args2[args2.index("--output") + 1] = str(tmp_path / "h")
    assert run(args2).returncode == 0
    assert filecmp.cmp(str(tmp_path / files[0]), str(tmp_path / files[0].replace("g_", "h_", 1)), shallow=False)


def test_generate_workload_sisi(tmp_path):
    r = run([os.path.join(BIN, "generate_dataset"), "--workload", "sisi_2p20", "--seed", "3",
             "--output", str(tmp_path / "w")])
    assert r.returncode == 0, r.stdout
    c = datasets.read_bin(str(tmp_path / "w_asc_sisi_2p20.bin"))
    assert list(c.sizes) == [1 << 20, 1 << 20] and c.universe == (1 << 24) - 1
    for i in range(2):
        s = c.set(i).astype(np.int64)
        assert np.all(np.diff(s) > 0) and s[-1] < (1 << 24)


def test_degree_cdf(tmp_path):
    src = os.path.join(GOLDEN, "random_k12.bin")
    r = run([os.path.join(BIN, "degree_cdf"), src, str(tmp_path / "s")])
    assert r.returncode == 0, r.stdout
    c = datasets.read_bin(src)
    deg = np.bincount(c.elements, minlength=c.universe + 1)
    read = lambda n: np.loadtxt(str(tmp_path / f"s_{n}.txt"), dtype=np.int64, ndmin=2)
    ds = read("degree_set")
    assert list(ds[:, 0]) == sorted(c.sizes) and sorted(ds[:, 1]) == list(range(1, c.k + 1))
    de = read("degree_element")
    assert de.shape[0] == deg.size and list(de[:, 0]) == sorted(deg)
    ec = read("element_cdf")
    assert ec[-1, 1] == int((deg.astype(np.int64) ** 2).sum())
    sc = read("set_cdf")
    assert sc[-1, 1] == int(sum(deg[c.set(i)].sum() for i in range(c.k)))


@pytest.fixture(scope="module")
def check_tune_store(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("ts") / "check_tune_store")
    r = run(["g++", "-O2", "-std=c++17", "-I" + HOST, "-I" + os.path.join(ROOT, "include"), "-o", exe,
             os.path.join(ROOT, "tests", "cpp", "check_tune_store.cpp")])
    assert r.returncode == 0, r.stderr
    return exe


def test_tune_store(check_tune_store, tmp_path):
    """The batched dispatcher's tuning store (host/tune_store.hpp): feature
    closed form, save/load round trip, exact / nearest / no lookup, and a
    damaged file whose bad records are skipped."""
    r = run([check_tune_store, str(tmp_path)])
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr


@pytest.mark.parametrize("name", ["random_k12", "random_k23"])
def test_tune_features_match_degree_cdf(check_tune_store, name, tmp_path):
    """The element-degree features the store keys on are degree_cdf's
    statistics: sum_e deg(e)^2 is the last value of its element CDF and the
    distinct elements are the elements of degree > 0."""
    src = os.path.join(GOLDEN, name + ".bin")
    r = run([check_tune_store, str(tmp_path), src])
    assert r.returncode == 0, r.stderr
    distinct, deg_mean, deg2, total = r.stdout.split()
    r = run([os.path.join(BIN, "degree_cdf"), src, str(tmp_path / "s")])
    assert r.returncode == 0, r.stdout
    ec = np.loadtxt(str(tmp_path / "s_element_cdf.txt"), dtype=np.int64, ndmin=2)
    de = np.loadtxt(str(tmp_path / "s_degree_element.txt"), dtype=np.int64, ndmin=2)
    assert int(distinct) == int((de[:, 0] > 0).sum())
    assert abs(float(deg2) * int(total) - ec[-1, 1]) < 1e-6 * ec[-1, 1]
    assert abs(float(deg_mean) * int(distinct) - int(total)) < 1e-9 * int(total)
