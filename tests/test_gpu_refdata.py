"""GPU tests of the reference's dataset formats: raw tuple files in and out
of HBM, the command-line driver on a data configuration file, and joins on
relations drawn by the restated reference generator."""
import os
import subprocess

import numpy as np
import pytest

import hjmi
import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "hashjoin-scm_amd", "bin", "hjmain")


def test_bin_round_trip(ctx, tmp_path):
    R, S = ctx.generate("zipf", 3000, 70000, 64, theta=1.2)
    for rel, name in ((R, "r.bin"), (S, "s.bin")):
        p = str(tmp_path / name)
        ctx.bin_write(p, rel)
        assert os.path.getsize(p) == rel.nbytes
        back = ctx.bin_read(p, rel.tuple_num, rel.tuple_size)
        assert np.array_equal(ctx.download(back), ctx.download(rel))
        ctx.free_relation(back)
    with pytest.raises(hjmi.HJError) as e:
        ctx.bin_read(str(tmp_path / "r.bin"), 3001, 64)  # one tuple more than the file holds
    assert e.value.code == hjmi.HJ_EINVAL
    with pytest.raises(hjmi.HJError):
        ctx.bin_read(str(tmp_path / "none.bin"), 1, 16)
    ctx.free_relation(R)
    ctx.free_relation(S)


def test_bin_read_large_file_in_chunks(ctx, tmp_path):
    """More than two 64 MiB staging buffers' worth (the double-buffered path)."""
    hR = np.random.default_rng(3).integers(0, 1 << 62, size=(160 << 20) // 8, dtype=np.uint64)
    p = str(tmp_path / "big.bin")
    hR.tofile(p)
    rel = ctx.bin_read(p, hR.nbytes // 16, 16)
    assert np.array_equal(ctx.download(rel).view(np.uint64), hR)
    ctx.free_relation(rel)


@pytest.mark.parametrize("kind,nr,ns,zf", [("dupfk", 1 << 12, 1 << 16, 0.0), ("zipf", 1 << 12, 1 << 16, 1.25),
                                           ("pkfk", 1 << 12, 1 << 16, 0.0)])
def test_cli_runs_reference_data_configuration(tmp_path, kind, nr, ns, zf):
    """Files in the reference's layout (exact reference order, written from
    the glibc-driven oracle) and a data configuration naming them: the
    driver's joins equal the oracle's aggregate."""
    hR, hS = O.ref_generate(kind, nr, ns, 16, zipf_factor=zf, exact=True, threads=4)
    (tmp_path / "uniform").mkdir()
    hR.tofile(tmp_path / "uniform" / "r.bin")
    hS.tofile(tmp_path / "uniform" / "s.bin")
    cfg = tmp_path / "w.cfg"
    cfg.write_text(f'r_path_suffix = "uniform/r.bin";\ns_path_suffix = "uniform/s.bin";\n'
                   f"r_tuple_num = {nr};\ns_tuple_num = {ns};\nr_tuple_size = 16;\ns_tuple_size = 16;\n"
                   f'theta = 0.0;\nmin_key = 1;\nmax_key = {nr};\nworkload_name = "mini_{kind}";\n')
    want = O.aggregate(hR, hS, 16)
    for algo in ("phj_radix_bc", "nphj_sc", "phj_radix_hr"):
        out = subprocess.run([CLI, f"--algo={algo}", f"--cfg={cfg}", f"--data-prefix={tmp_path}", "--trials=2"],
                             capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr
        lines = [l for l in out.stdout.splitlines() if l.startswith("matched_cnt:")]
        assert len(lines) == 2
        for l in lines:
            f = l.split()
            assert (int(f[1]), int(f[3])) == (want["matched_cnt"], want["check_sum"]), (algo, l)
        assert f"workload: mini_{kind}" in out.stdout


def test_reference_generator_relations_join_like_oracle(ctx):
    """dupfk at a mid size: duplicates on both sides, every all-pairs join
    equals the oracle's per-key aggregation."""
    sp = hjmi.RefSpec(hjmi.HJ_REF_DUPFK, 16, 1 << 16, 1 << 20, 0.0, 12345, 0)
    hR, hS = hjmi.ref_generate(sp)
    want = O.aggregate(hR, hS, 16)
    R, S = ctx.upload(hR, 16), ctx.upload(hS, 16)
    for algo in ("phj_radix_bc", "phj_radix_sc", "phj_radix_hr", "phj_radix_hro", "nphj_sc"):
        got = ctx.join(algo, R, S)
        assert (got["matched_cnt"], got["check_sum"]) == (want["matched_cnt"], want["check_sum"]), algo
    ctx.free_relation(R)
    ctx.free_relation(S)
