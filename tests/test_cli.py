"""The rtxrmq driver: the reference's command-line contract (src/main.cu, src/tools.h)."""

import csv
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gpu-rmq_amd", "rtxrmq")
XXX_HEADER = ("dev,alg,nt,reps,n,bs,nb,q,lr,GPU_BSIZE,CG_GROUP_SIZE,CG_MEM_ALIGNMENT,timestamp,"
              "scan_threshold,XXX_CG_SIZE_LOG,XXX_CG_AMOUNT_LOG,t,q/s,ns/q,construction,outbuffer,"
              "tempbuffer,freeGPUMemCorrect,freeGPUMem,checkResult")


def run(*args, timeout=300):
    return subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, timeout=timeout)


@pytest.mark.parametrize("argv", [
    (),                      # too few arguments -> help, EXIT_FAILURE (src/tools.h:93-96)
    (1000, 10, 0, 16),       # lr = 0 -> help (src/tools.h:107-110)
    (1000, 10, 2000, 16),    # lr > n (src/tools.h:115-118)
    (1000, 10, -1, 17),      # an algorithm this driver does not provide
    (1000, 10, -1, 1),       # HRMQ: an external library, not part of this driver
    (1000, 10, -1, 101),
    (1000, 10, -1, 16, "--nb", 0),  # the reference divides by it (src/tools.h:174)
])
def test_argument_errors_exit_nonzero(argv):
    p = run(*argv)
    assert p.returncode == 1, (argv, p.stdout, p.stderr)


@pytest.mark.gpu
@pytest.mark.parametrize("alg", [16, 20, 24, 116, 120, 124, 2, 102])
def test_check_passes_and_csv_schema(alg, tmp_path, device):
    csv_file = tmp_path / "data.csv"
    p = run(200003, 50000, -6, alg, "--seed", 27722, "--reps", 2, "--check", "--log_bs", 10,
            f"--save-time={csv_file}")
    assert p.returncode == 0, p.stdout + p.stderr
    assert "pass" in p.stdout
    lines = csv_file.read_text().splitlines()
    rows = list(csv.reader(lines))
    if alg % 100 in (16, 20, 24):
        assert lines[0] == XXX_HEADER
        assert len(rows[1]) == len(rows[0])
        assert rows[1][-1] == "1" and rows[1][13] == str(1 << 10)
    else:
        assert len(rows[1]) == len(rows[0])
        assert rows[1][-1] == "1"


@pytest.mark.gpu
def test_threshold_error_completes_csv_line(tmp_path, device):
    csv_file = tmp_path / "data.csv"
    p = run(100000, 1000, -1, 16, "--log_bs", 5, "--seed", 1, f"--save-time={csv_file}")
    assert p.returncode == 1
    assert "must be at least 2x reduction_factor (32)" in p.stderr
    assert csv_file.read_text().endswith("\n")


@pytest.mark.gpu
def test_rand_trivial_check_and_paper_constants(tmp_path, device):
    p = run(1 << 22, 1 << 20, -2, 16, "--seed", 913, "--reps", 1, "--randTrivialCheck",
            "--cg_size_log", 4, "--cg_amount_log", 1, "--log_bs", 6)
    assert p.returncode == 0 and "pass" in p.stdout, p.stdout + p.stderr
    p = run(1 << 22, 1 << 20, -3, 20, "--seed", 913, "--reps", 1, "--trivialCheck",
            "--cg_amount_log", 1, "--log_bs", 5)
    assert p.returncode == 0 and "pass" in p.stdout, p.stdout + p.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("alg", [16, 20, 24, 116])
def test_reference_default_case(alg, tmp_path, device):
    """The reference's own small case with its defaults: n = q = 2^20, uniform lengths (lr = -1),
    no --log_bs (t = 2^15, so long ranges end on a top level of thousands of entries), checked
    on 2^15 random queries against GPU BASE."""
    csv_file = tmp_path / "data.csv"
    p = run(1 << 20, 1 << 20, -1, alg, "--seed", 27722, "--reps", 2, "--randTrivialCheck",
            f"--save-time={csv_file}")
    assert p.returncode == 0 and "pass" in p.stdout, p.stdout + p.stderr
    rows = list(csv.reader(csv_file.read_text().splitlines()))
    assert rows[1][13] == str(1 << 15) and rows[1][-1] == "5"  # pass (1) + 4 for the random check


@pytest.mark.gpu
@pytest.mark.parametrize("alg", [20, 120, 116, 124, 102])
def test_long_positions_check(alg, tmp_path, device):
    """--long: the reference's IS_LONG build (64-bit {l, r} pairs and indices) on an array below
    2^31, checked bit-exactly against 64-bit GPU BASE."""
    p = run(300007, 40000, -6, alg, "--seed", 5, "--reps", 1, "--check", "--long", "--log_bs", 10)
    assert p.returncode == 0 and "pass" in p.stdout, p.stdout + p.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("alg", [20, 120, 116])
def test_pack_descents_check(alg, device):
    """--pack_descents (the lane solver with its descents packed) passes the bit-exact --check
    against GPU BASE on lognormal-medium and mixed lengths."""
    for lr in (-2, -6):
        p = run(1 << 20, 100000, lr, alg, "--seed", 41, "--reps", 1, "--check", "--log_bs", 10,
                "--cg_amount_log", 0 if alg % 100 == 20 else 2, "--pack_descents")
        assert p.returncode == 0 and "pass" in p.stdout, (lr, p.stdout + p.stderr)


@pytest.mark.gpu
def test_long_array_above_2e31(tmp_path, device):
    """n = 2^31 + 7 turns the 64-bit form on by itself; a random sample is checked. Short ranges
    (lr = -3): the check is GPU BASE, one lane scanning each range, so long ranges would take
    minutes at this n (the long ranges of this form are checked in test_long.py)."""
    p = run((1 << 31) + 7, 1 << 20, -3, 120, "--seed", 3, "--reps", 1, "--randTrivialCheck",
            "--log_bs", 13, "--cg_amount_log", 1, timeout=300)
    assert p.returncode == 0 and "pass" in p.stdout, p.stdout + p.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("alg", [0, 100])
def test_cpu_base_check_and_csv(alg, tmp_path, device):
    """Algs 0 / 100: the CPU BASE scan (OpenMP over queries, --nt threads) passes the bit-exact
    --check against GPU BASE; its CSV row names the host and a block size of 0, as the
    reference's does (src/tools.h:359-365)."""
    csv_file = tmp_path / "data.csv"
    p = run(200003, 20000, -6, alg, "--seed", 27722, "--reps", 2, "--nt", 4, "--check",
            f"--save-time={csv_file}")
    assert p.returncode == 0 and "pass" in p.stdout, p.stdout + p.stderr
    assert ("Computing RMQs index ([CPU] BASE IDX, nt= 4" if alg else "Computing RMQs ([CPU] BASE , nt= 4") in p.stdout
    rows = list(csv.reader(csv_file.read_text().splitlines()))
    assert rows[1][0].startswith("CPU ") and rows[1][9] == "0" and rows[1][-1] == "1"
    assert rows[1][1] == ("[CPU] BASE IDX" if alg else "[CPU] BASE") and len(rows[1]) == len(rows[0])


@pytest.mark.gpu
def test_nb_sets_the_scan_threshold(tmp_path, device):
    """--nb B sets t = bit_ceil(n / B) (src/tools.h:168-175): n = 200003, B = 100 gives 2048,
    written to the CSV's scan_threshold column."""
    csv_file = tmp_path / "data.csv"
    p = run(200003, 20000, -6, 16, "--seed", 3, "--reps", 1, "--nb", 100, "--randTrivialCheck",
            f"--save-time={csv_file}")
    assert p.returncode == 0 and "pass" in p.stdout, p.stdout + p.stderr
    assert "bs   = 2048" in p.stdout
    rows = list(csv.reader(csv_file.read_text().splitlines()))
    assert rows[1][13] == "2048"


@pytest.mark.gpu
@pytest.mark.parametrize("alg,mode", [(16, "--check"), (116, "--check"), (120, "--randTrivialCheck"),
                                      (20, "--trivialCheck")])
def test_save_input_data_cache(alg, mode, tmp_path, device):
    """--save-input-data writes the array, the queries and the expected results under
    $RMQ_AUX_DIR with the reference's names and little-endian layouts (src/tools.h:526-617,
    src/cuda_methods.cuh:175-210: results named by the checked count, values only in value
    mode); a second run reads all of them back and passes."""
    n, q, lr, seed = 100003, 40000, -6, 99
    k = {"--check": q, "--randTrivialCheck": min(q, 1 << 15), "--trivialCheck": min(q, 1 << 10)}[mode]
    triv = int(mode == "--trivialCheck")
    rand = int(mode == "--randTrivialCheck")
    env = dict(os.environ, RMQ_AUX_DIR=str(tmp_path))
    args = [CLI, n, q, lr, alg, "--seed", seed, "--reps", 1, "--log_bs", 10, mode, "--save-input-data"]
    p = subprocess.run(list(map(str, args)), capture_output=True, text=True, timeout=300, env=env)
    assert p.returncode == 0 and "pass" in p.stdout, p.stdout + p.stderr
    tail = f"_lr-{lr}_seed-{seed}_trivCheck-{triv}_randTrivCheck-{rand}.bin"
    files = {f"array_n-{n}_seed-{seed}.bin": 4 * n, f"queries_n-{n}_q-{q}{tail}": 8 * q,
             f"results_index_n-{n}_q-{k}{tail}": 4 * k}
    if alg < 100:
        files[f"results_value_n-{n}_q-{k}{tail}"] = 4 * k
    assert sorted(os.listdir(tmp_path)) == sorted(files)
    for name, size in files.items():
        assert os.path.getsize(tmp_path / name) == size, name
    # the cached array is the generator's, byte for byte
    import numpy as np

    import rmq_amd

    assert np.array_equal(np.fromfile(tmp_path / f"array_n-{n}_seed-{seed}.bin", dtype=np.float32).view(np.uint32),
                          rmq_amd.gen_array(n, seed).view(np.uint32))
    p = subprocess.run(list(map(str, args)), capture_output=True, text=True, timeout=300, env=env)
    assert p.returncode == 0 and "pass" in p.stdout, p.stdout + p.stderr
    assert p.stdout.count("Reading data in file") == len(files), p.stdout
    assert "Writing data in file" not in p.stdout
