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
    (1000, 10, -1, 0),       # CPU BASE lives in the test oracle, not in the driver
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
