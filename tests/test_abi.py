"""The C-ABI library loads and exports every entry point include/actflow.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import os
import re
import subprocess

from conftest import ROOT


def declared(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(af_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_symbols_exported(af):
    lib = af.lib()
    names = declared(os.path.join(ROOT, "include", "actflow.h"))
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_no_oracle_linkage():
    """The product library neither links nor names the oracle."""
    so = os.path.join(ROOT, "2d_active_matter_simulation_amd", "libactflow.so")
    out = subprocess.run(["nm", "-D", so], capture_output=True, text=True).stdout
    assert "oracle_" not in out
    ldd = subprocess.run(["ldd", so], capture_output=True, text=True).stdout
    assert "liboracle" not in ldd
    assert "amdhip64" in ldd


def test_status_strings(af):
    lib = af.lib()
    for s in range(7):
        assert lib.af_status_str(s)


def test_oracle_exports(orc):
    lib = orc.lib()
    for n in ("oracle_create", "oracle_step", "oracle_rhs", "oracle_rfft2", "oracle_irfft2", "oracle_fft1"):
        assert hasattr(lib, n)


def test_params_mapping(af):
    p = af.params(256, alpha=-0.8)
    assert (p.c0, p.c2, p.c4) == (-1.0 + 0.8, 2.0, -1.0)
    q = af.params(256, alpha=-0.8, gamma0=-0.045, gamma2=2.5e-5)
    assert (q.c0, q.c2, q.c4) == (0.8, 0.045, -2.5e-5)


def test_create_rejects_bad_sizes(af):
    """Argument checks run before any device call."""
    for nx, ny, prec, dea in ((48, 64, 64, 0), (64, 8192, 64, 0), (64, 64, 16, 0), (64, 64, 64, 2)):
        p = af.params(nx, ny, precision=prec, dealias=dea)
        try:
            af.Run(p)
        except af.ActFlowError as e:
            assert e.status in (af.AF_EINVAL, af.AF_EUNSUPPORTED)
        else:
            raise AssertionError("accepted bad parameters")


def test_tuned_tables_name_instantiated_schedules(af):
    """Every line of the shipped tuned tables (schedules.txt for single runs,
    schedules_batched.txt for the engine) names a schedule the library has
    kernels for, so no tuned size silently falls back to the default."""
    pkg = os.path.join(ROOT, "2d_active_matter_simulation_amd")
    for name in ("schedules.txt", "schedules_batched.txt"):
        path = os.path.join(pkg, name)
        assert os.path.exists(path), name
        for line in open(path):
            body = line.split("#")[0].split()
            if not body:
                continue
            v = [int(x) for x in body]
            prec, nx, ny, sched = v[0], v[1], v[2], tuple(v[3:])
            if len(sched) == 7:
                sched = sched + (sched[1],)
            if len(sched) == 8:
                sched = sched + (0,)
            assert sched in af.schedules(prec, nx, ny), (name, line)


def test_schedule_struct_fields(af):
    """af_schedule carries the column-block width (divides the brick width)
    and the option bits (LDS double buffering; streaming stores, FP32 only;
    paired rows, with an even number of rows per row block; half-brick row
    blocks, FP32 only, with an even brick height; P, Q in taller and U, V, W
    in shorter bricks, not with paired rows)."""
    for prec in (32, 64):
        for sc in af.schedules(prec, 1024):
            by, bx, rr, rg, cr, gf, gi, cw, opts = sc
            assert bx % cw == 0 and rr in (4, 8, 16) and cr in (4, 8, 16)
            assert 0 <= opts < 512 and (prec == 32 or opts & ~32 in (0, 1))
            rows = by // 2 if opts & 64 else by
            assert not opts & 64 or by % 2 == 0
            assert not opts & 32 or rows % 2 == 0
            assert not (opts & 128 and opts & 32)
            assert not opts & 256 or (opts & 64 and not opts & 32)
    assert any(sc[8] & 6 for sc in af.schedules(32, 1024))
    assert any(sc[8] & 32 for sc in af.schedules(32, 1024)) and any(sc[8] & 32 for sc in af.schedules(64, 1024))
    assert any(sc[8] & 64 for sc in af.schedules(32, 4096)) and any(sc[8] & 128 for sc in af.schedules(32, 4096))
    assert any(sc[8] & 256 for sc in af.schedules(32, 1024))


def test_autotune_merge_keeps_untuned_lines():
    """Re-tuning some sizes rewrites only their lines of the table."""
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "autotune", os.path.join(ROOT, "2d_active_matter_simulation_amd", "autotune.py"))
    at = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(at)
    old = "# header\n32 1024 1024 4 2 4 3 4 2 3 2 0   # 0.1\n64 1024 1024 4 2 4 3 4 2 3 2 0   # 0.2\n"
    new = at.merge_table(["# header", "32 1024 1024 4 2 4 3 4 2 1 2 15   # 0.09"], old)
    rows = [ln for ln in new.splitlines() if not ln.startswith("#")]
    assert rows == ["32 1024 1024 4 2 4 3 4 2 1 2 15   # 0.09", "64 1024 1024 4 2 4 3 4 2 3 2 0   # 0.2"]


def build_c_example(out_dir):
    """tests/c_abi/step_loop.c: a plain C99 caller of the ABI (strict flags,
    warnings as errors), linked against the library."""
    exe = os.path.join(str(out_dir), "step_loop")
    lib_dir = os.path.join(ROOT, "2d_active_matter_simulation_amd")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror",
                    "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "c_abi", "step_loop.c"),
                    "-L", lib_dir, "-lactflow", f"-Wl,-rpath,{lib_dir}", "-o", exe], check=True)
    return exe


def test_header_is_strict_c99_and_links(tmp_path):
    """The header is the boundary for C callers: a C99 program that drives
    the reference's loop through it compiles with -pedantic -Werror and
    links (it runs in test_gpu_c_abi.py)."""
    assert os.path.exists(build_c_example(tmp_path))
