"""CPU checks of the host side: the C-ABI library loads and exports every
entry point include/kemorin_sht.h declares, argument validation fails with
SHT_ERR_ARG before touching a device, and the Python mirror's index maps and
seeded spectra follow the reference's single-rank conventions."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import kemorin_mhd_amd as K
from kemorin_mhd_amd import sht
from tests.golden.make_sph_files import paths as sph_fixture_paths

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "kemorin_sht.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sht_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("sht_plan_create", "sht_plan_destroy", "sht_backward", "sht_forward",
                 "sht_legendre_backward", "sht_legendre_forward", "sht_fft_backward",
                 "sht_fft_forward", "sht_exchange_rtm_to_rtp", "sht_exchange_rtp_to_rtm",
                 "sht_exchange_gather_rev", "sht_get_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(sht.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", sht.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (sht_\w+)", out))
    assert set(declared_functions()) <= exported


def test_version_string():
    v = sht.lib().sht_version().decode()
    assert "gfx950" in v or len(v) > 0


def _desc(**kw):
    r = np.linspace(0.5, 1.5, kw.get("nr", 4))
    d = sht._Desc(kw.get("L", 7), kw.get("nr", 4), kw.get("nth", 12), kw.get("nphi", 16),
                  r.ctypes.data_as(sht._dp), kw.get("ncomp", 9), kw.get("prec", 0), 0, 0)
    return d, r


@pytest.mark.parametrize("bad", [dict(L=0), dict(nth=5), dict(nphi=14), dict(nphi=17),
                                 dict(ncomp=0), dict(ncomp=65), dict(prec=7)])
def test_plan_create_rejects_bad_descriptors(bad):
    d, r = _desc(**bad)
    h = ctypes.c_void_p()
    rc = sht.lib().sht_plan_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == 1 and not h.value
    assert "sht_plan_create" in sht.last_error()


def test_batch_create_rejects_bad_arguments():
    h = ctypes.c_void_p()
    L = sht.lib()
    assert L.sht_batch_create(None, 1, ctypes.byref(h)) == 1 and not h.value
    calls = (sht._ShellCall * 1)()
    assert L.sht_batch_create(calls, 0, ctypes.byref(h)) == 1
    assert L.sht_batch_create(calls, 1, ctypes.byref(h)) == 1  # no plan
    assert "shell 0 has no plan" in sht.last_error()
    assert L.sht_batch_launch(None, None) == 1
    assert L.sht_batch_destroy(None) == 0


def test_plan_create_rejects_null_and_nonpositive_radius():
    h = ctypes.c_void_p()
    assert sht.lib().sht_plan_create(None, ctypes.byref(h)) == 1
    d, r = _desc()
    r[2] = 0.0
    assert sht.lib().sht_plan_create(ctypes.byref(d), ctypes.byref(h)) == 1
    assert "radius" in sht.last_error()


def test_variants_and_autotune_reject_null_plan():
    L = K.lib()
    v = ctypes.c_int()
    assert L.sht_plan_set_variant(None, b"split_fwd_mfma", 32) == 1
    assert L.sht_plan_get_variant(None, b"fft_form", ctypes.byref(v)) == 1
    assert L.sht_plan_autotune(None, 2, 2, 3, 0, None, None) == 1
    assert "null" in K.last_error()


def test_calls_on_null_plan_fail_cleanly():
    L = sht.lib()
    assert L.sht_backward(None, 1, 0, None, None) != 0
    assert L.sht_pole_backward(None, 1, 0, None, None, 1) != 0
    assert L.sht_backward_with_poles(None, 1, 0, None, None, None, 0) != 0
    assert L.sht_nonlinear_rtp(None, None, 3, None, 3, None) != 0
    assert L.sht_plan_destroy(None) == 0 or True  # destroying NULL is a no-op


def test_shell_config_sizes():
    c = K.ShellConfig(511, 257, 768, 1536)
    assert c.jmax == 512 ** 2
    assert c.n_rj(8) == 8 * 512 ** 2 * 257
    assert c.n_rtp(9) == 9 * 257 * 768 * 1536
    assert c.n_rtm(9) == 9 * 768 * 257 * 1023


def test_rlm_order():
    """single-rank rlm order: m from -L to L, then l = |m| .. L
    (CG/set_sph_tranform_ordering.f90:69-206)."""
    l, m = K.rlm_l_m(3)
    assert len(l) == 16
    assert list(zip(l[:4], m[:4])) == [(3, -3), (2, -2), (3, -2), (1, -1)]
    assert (l[m == 0] == np.arange(4)).all()
    assert np.all(np.diff(m) >= 0)


def test_synthetic_spectra_deterministic_and_shaped():
    c = K.ShellConfig(15, 5, 24, 40)
    a = K.synthetic_spectra(c, 2, 1, seed=3)
    b = K.synthetic_spectra(c, 2, 1, seed=3)
    assert np.array_equal(a, b)
    s = a.reshape(c.nr, c.jmax, 7)
    assert np.all(s[0] == 0) and np.all(np.abs(s[-1]) < 1e-15)  # sin profile vanishes at ri, ro
    l, _ = K.rlm_l_m(c.L)
    rms = [np.sqrt(np.mean(s[2][l == ll] ** 2)) for ll in (1, 15)]
    assert rms[0] > 10 * rms[1]  # (l+1)^-1.5 decay


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stdout


AMDFLANG = "/opt/rocm/bin/amdflang"


_FORTRAN_DRIVER = """
      program drv
      use iso_c_binding
      use kemorin_sht_c
      implicit none
      type(sht_desc) :: d
      type(sht_nonlinear_desc) :: nl
      type(c_ptr) :: plan, files
      real(c_double), target :: r(4)
      real(c_double) :: rr(16)
      integer(c_int) :: rc, rc2, icase, dims(5), items(200)
      integer(c_long_long) :: nitem
      r = (/0.5d0, 0.8d0, 1.1d0, 1.5d0/)
      do icase = 0, 10
        d%l_truncation = 3
        d%nidx_r = 4
        d%nidx_theta = 12
        d%nidx_phi = 16
        d%radius = c_loc(r)
        d%max_ncomp = 9
        d%precision = SHT_PREC_FP64
        d%n_slices = 0
        d%flags = 0
        r(2) = 0.8d0
        select case (icase)
        case (1)
          d%l_truncation = 0
        case (2)
          d%nidx_r = 0
        case (3)
          d%nidx_theta = 3
        case (4)
          d%nidx_phi = 15
        case (5)
          d%radius = c_null_ptr
        case (6)
          d%max_ncomp = 65
        case (7)
          d%precision = 7
        case (8)
          d%precision = SHT_PREC_SPLIT_INT8
          d%n_slices = 3
        case (9)
          d%flags = 5
        case (10)
          r(2) = -1.0d0
        end select
        rc = sht_plan_create(d, plan)
        write(*, '(a,i0,a,i0,2a)') 'case ', icase, ' rc=', rc, ' ', trim(last_error())
      end do
      nl%i_velo = 0
      nl%i_coriolis = 0
      rc = sht_nonlinear_rtp(c_null_ptr, nl, 3_c_int, c_null_ptr,       &
     &                       3_c_int, c_null_ptr)
      write(*, '(a,i0)') 'nl=', rc
      rc = sht_sph_files_read('@RTP@' // c_null_char,                   &
     &    '@RTM@' // c_null_char, '@RLM@' // c_null_char,               &
     &    '@RJ@' // c_null_char, files)
      write(*, '(a,i0)') 'files=', rc
      rc = sht_sph_files_check(files)
      rc2 = sht_sph_files_shape(files, dims, rr)
      nitem = sht_sph_files_table(files, 'rj_import' // c_null_char,   &
     &                            items, 200_c_long_long)
      write(*, '(a,i0,a,i0,a,5(1x,i0),a,i0,a,i0,a,i0)') 'check=', rc,      &
     &    ' shape=', rc2, ' dims', dims, ' items=', nitem,              &
     &    ' first=', items(1), ' last=', items(nitem)
      call sht_sph_files_destroy(files)
      contains
      function last_error() result(msg)
      character(len = 256) :: msg
      character(kind = c_char) :: buf(256)
      integer(c_int) :: i, irc
      buf = c_null_char
      irc = sht_get_last_error(buf, 256_c_int)
      msg = ' '
      do i = 1, 256
        if (buf(i) .eq. c_null_char) exit
        msg(i:i) = buf(i)
      end do
      end function last_error
      end program drv
"""

# field -> (case number, text the message must carry: the field and its value)
_FORTRAN_CASES = {
    "l_truncation": (1, "l_truncation = 0"), "nidx_r": (2, "nidx_r = 0"),
    "nidx_theta": (3, "nidx_theta = 3"), "nidx_phi": (4, "nidx_phi = 15"),
    "radius": (5, "radius: null"), "max_ncomp": (6, "max_ncomp = 65"),
    "precision": (7, "precision = 7"), "n_slices": (8, "n_slices = 3"),
    "flags": (9, "flags = 5"), "radius values": (10, "radius[1] must be positive"),
}


@pytest.mark.skipif(not os.path.exists(AMDFLANG), reason="no Fortran compiler")
def test_fortran_binding_lays_out_every_descriptor_field(tmp_path):
    """The ISO_C_BINDING module compiles and links against the library.  A
    Fortran caller breaks one sht_desc field at a time and gets SHT_ERR_ARG
    with the message naming that field and the value it set, which pins
    every field's offset in the bind(C) type; the unbroken descriptor passes
    validation (and then finds no device here: SHT_ERR_HIP)."""
    mod = os.path.join(ROOT, "kemorin_mhd_amd", "fortran", "kemorin_sht_c.f90")
    prog = tmp_path / "drv.f90"
    drv = _FORTRAN_DRIVER
    for tag, path in zip(("@RTP@", "@RTM@", "@RLM@", "@RJ@"), sph_fixture_paths()):
        drv = drv.replace(tag, path)
    prog.write_text(drv)
    exe = tmp_path / "drv"
    libdir = os.path.dirname(sht.LIB_PATH)
    r = subprocess.run([AMDFLANG, "-J", str(tmp_path), mod, str(prog), "-o", str(exe),
                        f"-L{libdir}", "-lkemorin_sht", f"-Wl,-rpath,{libdir}"],
                       capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120,
                         env={**os.environ, "HIP_VISIBLE_DEVICES": ""})
    lines = {int(ln.split()[1]): ln for ln in out.stdout.splitlines() if ln.startswith("case ")}
    assert len(lines) == 11, out.stdout + out.stderr
    assert " rc=4 " in lines[0] or lines[0].rstrip().endswith("rc=4"), lines[0]
    for field, (case, text) in _FORTRAN_CASES.items():
        assert " rc=1 " in lines[case], (field, lines[case])
        assert text in lines[case], (field, lines[case])
    assert "nl=1" in out.stdout, out.stdout + out.stderr
    # the grid / mode files through the binding (no device needed)
    assert "files=0" in out.stdout, out.stdout + out.stderr
    assert "check=0 shape=0 dims 5 3 9 18 108 items=108 first=1 last=108" in out.stdout, out.stdout


# --- the split GEMMs' hand-counted waits (checker of the generated code) ---
_K = "_ZN4ksht9k_leg_toyEv:\n"


def _waits(body):
    from kemorin_mhd_amd.isa_waits import check
    return check(_K + body + "\ts_endpgm\n", "k_leg_")["_ZN4ksht9k_leg_toyEv"]


def test_wait_check_accepts_counted_waits():
    body = ("\tglobal_load_dwordx4 v[2:5], v1, s[0:1]\n"
            "\tglobal_load_lds_dwordx4 v1, s[2:3]\n"
            "\ts_waitcnt vmcnt(1)\n"
            "\tv_mfma_i32_16x16x64_i8 v[6:9], v[2:5], v[10:13], 0\n")
    assert _waits(body) == []


def test_wait_check_flags_use_and_copy_before_the_wait():
    use = ("\tglobal_load_dwordx4 v[2:5], v1, s[0:1]\n"
           "\tglobal_load_lds_dwordx4 v1, s[2:3]\n"
           "\ts_waitcnt vmcnt(2)\n"
           "\tv_mov_b64_e32 v[20:21], v[2:3]\n")
    assert len(_waits(use)) == 1
    reuse = ("\tglobal_load_dwordx4 v[2:5], v1, s[0:1]\n"
             "\tv_mov_b32_e32 v4, 0\n"
             "\ts_waitcnt vmcnt(0)\n")
    assert len(_waits(reuse)) == 1


def test_wait_check_load_order():
    # a later load into the registers of one still in flight lands last
    # (loads return in issue order): accepted; using them as an address is not
    waw = ("\tglobal_load_dwordx4 v[2:5], v1, s[0:1]\n"
           "\tglobal_load_dwordx4 v[4:7], v8, s[0:1]\n"
           "\ts_waitcnt vmcnt(0)\n")
    assert _waits(waw) == []
    addr = ("\tglobal_load_dwordx2 v[2:3], v1, s[0:1]\n"
            "\tglobal_load_dwordx4 v[4:7], v2, s[0:1]\n"
            "\ts_waitcnt vmcnt(0)\n")
    assert len(_waits(addr)) == 1


def test_wait_check_follows_every_path():
    # the copy after the branch is skipped on one path, so vmcnt(1) does
    # not retire the register load there
    body = ("\tglobal_load_dwordx4 v[2:5], v1, s[0:1]\n"
            "\ts_cbranch_scc1 .LBB0_2\n"
            "\tglobal_load_lds_dwordx4 v1, s[2:3]\n"
            ".LBB0_2:\n"
            "\ts_waitcnt vmcnt(1)\n"
            "\tv_add_u32_e32 v30, v2, v31\n")
    assert len(_waits(body)) == 1
    assert _waits(body.replace("vmcnt(1)", "vmcnt(0)")) == []


def test_wait_check_flags_scratch():
    assert any("scratch" in p for p in _waits("\tscratch_store_dword off, v4, off\n"))


def test_wait_check_flags_missing_wait_states():
    m0 = "\ts_mov_b32 m0, s4\n\tglobal_load_lds_dwordx4 v1, s[2:3]\n\ts_waitcnt vmcnt(0)\n"
    assert any("M0" in p for p in _waits(m0))
    assert _waits(m0.replace("m0, s4\n", "m0, s4\n\ts_nop 0\n")) == []
    sg = "\tv_readfirstlane_b32 s0, v7\n\tglobal_load_lds_dword v1, s[0:1]\n\ts_waitcnt vmcnt(0)\n"
    assert any("SGPR" in p for p in _waits(sg))
    assert any("SGPR" in p for p in _waits(sg.replace("v7\n", "v7\n\ts_nop 3\n")))
    assert _waits(sg.replace("v7\n", "v7\n\ts_nop 4\n")) == []
    # across a branch target: the write sits in a predecessor block
    br = ("\tv_readfirstlane_b32 s1, v7\n\ts_cbranch_scc1 .LBB0_1\n\ts_nop 7\n"
          ".LBB0_1:\n\tglobal_load_dwordx4 v[2:5], v1, s[0:1]\n\ts_waitcnt vmcnt(0)\n")
    assert any("SGPR" in p for p in _waits(br))


def test_built_libraries_are_current():
    """The shared libraries that travel with the tree (the HIP library and the
    oracle) are newer than every source they are built from."""
    csrc = os.path.join(ROOT, "kemorin_mhd_amd", "csrc")
    lib = os.path.join(ROOT, "kemorin_mhd_amd", "libkemorin_sht.so")
    srcs = [os.path.join(csrc, f) for f in os.listdir(csrc)] + [
        os.path.join(ROOT, "include", "kemorin_sht.h")]
    orc = os.path.join(ROOT, "oracle", "liboracle.so")
    for so, deps in ((lib, srcs), (orc, [os.path.join(ROOT, "oracle", "sht_oracle.c"),
                                         os.path.join(ROOT, "oracle", "sht_oracle.h")])):
        if not os.path.exists(so):
            pytest.skip(f"{so} not built")
        stale = [d for d in deps if os.path.getmtime(d) > os.path.getmtime(so)]
        assert not stale, f"{os.path.basename(so)} is older than {stale}"
