"""The C++ facade's .ini reader and parameter setup (parallel-advection_amd/
cpp/bkma/config.hpp) against the reference's own INI reader compiled from its
sources into oracle/_ref (oracle/ref.mk), on the reference's .ini files and on
files that exercise its parsing rules."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import bkma_amd as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref_ini.so")
REF_SRC = "/root/reference/src"

SAMPLES = {
    "comments_and_quirks.ini": (
        "; leading comment\n# hash comment\n[Problem]\nN0 = 0x10\nn1 = 010 ; octal with inline comment\n"
        "n2 = 12 # hash is not a comment here\nmaxIter=7;no space so kept\ndt = 0.001\n"
        "minRealVx = -1e-3\n[impl]\nkernelImpl = NDRange   \ninplace = true # trailing\n"
        "[optimization]\ngpu = off\npref_wg_size = 256\n  continued value\n"
        "seq_size0\n[broken\n[io]\noutputSolution = yes\n"),
    "long_lines.ini": (
        "[problem]\n" + "n0 = 5 " + "x" * 250 + "\n" + "[" + "s" * 70 + "]\n" + "k" * 60 + " = 3\n"
        "[problem]\ndt = 2.5e-4\r\nn1 = 77\r\n"),
    "empty_values.ini": "[problem]\nn0 =\ndt = abc\n[impl]\ninplace =\n[optimization]\ngpu = 1\n",
}

KEYS = [("problem", k) for k in ("n0", "n1", "n2", "maxIter", "dt", "minRealX", "maxRealX", "minRealVx",
                                 "maxRealVx", "length", "channel_in", "k", "total_batch_size",
                                 "batch_size_n2")] + \
       [("impl", "kernelImpl"), ("impl", "inplace"), ("optimization", "gpu"),
        ("optimization", "pref_wg_size"), ("optimization", "seq_size0"), ("optimization", "seq_size2"),
        ("io", "outputSolution"), ("s" * 49, "k" * 60), ("PROBLEM", "N0"), ("nosuch", "key")]


@pytest.fixture(scope="module")
def readers(tmp_path_factory):
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref/libref_ini.so not built (needs the reference tree)")
    d = tmp_path_factory.mktemp("ini")
    so = str(d / "libfac_ini.so")
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-o", so,
                    os.path.join(ROOT, "tests", "cpp", "facade_ini_shim.cpp"),
                    "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "parallel-advection_amd", "cpp"),
                    "-L" + os.path.dirname(B.LIB_PATH), "-Wl,-rpath," + os.path.dirname(B.LIB_PATH), "-lbkma_amd"],
                   check=True)
    libs = {}
    for tag, path in (("ref", REF_LIB), ("fac", so)):
        L = ctypes.CDLL(path)
        vp, cp = ctypes.c_void_p, ctypes.c_char_p
        getattr(L, f"{tag}_ini_open").restype = vp
        getattr(L, f"{tag}_ini_open").argtypes = [cp]
        getattr(L, f"{tag}_ini_close").argtypes = [vp]
        getattr(L, f"{tag}_ini_error").argtypes = [vp]
        getattr(L, f"{tag}_ini_get_integer").restype = ctypes.c_long
        getattr(L, f"{tag}_ini_get_integer").argtypes = [vp, cp, cp, ctypes.c_long]
        getattr(L, f"{tag}_ini_get_float").restype = ctypes.c_double
        getattr(L, f"{tag}_ini_get_float").argtypes = [vp, cp, cp, ctypes.c_float]
        getattr(L, f"{tag}_ini_get_bool").argtypes = [vp, cp, cp, ctypes.c_int]
        getattr(L, f"{tag}_ini_get_string").argtypes = [vp, cp, cp, cp, cp, ctypes.c_int]
        libs[tag] = L
    files = {}
    for name, text in SAMPLES.items():
        p = d / name
        p.write_bytes(text.encode())
        files[name] = str(p)
    for name in ("advection.ini", "conv1d.ini"):
        src = os.path.join(REF_SRC, name)
        if os.path.exists(src):
            files["ref_" + name] = src
    files["missing.ini"] = str(d / "does_not_exist.ini")
    return libs, files


def _read_all(L, tag, path):
    h = getattr(L, f"{tag}_ini_open")(path.encode())
    out = {"error": getattr(L, f"{tag}_ini_error")(h)}
    buf = ctypes.create_string_buffer(512)
    for s, k in KEYS:
        sb, kb = s.encode(), k.encode()
        out[(s, k, "int")] = getattr(L, f"{tag}_ini_get_integer")(h, sb, kb, -12345)
        out[(s, k, "float")] = getattr(L, f"{tag}_ini_get_float")(h, sb, kb, ctypes.c_float(0.25))
        out[(s, k, "bool_t")] = getattr(L, f"{tag}_ini_get_bool")(h, sb, kb, 1)
        out[(s, k, "bool_f")] = getattr(L, f"{tag}_ini_get_bool")(h, sb, kb, 0)
        getattr(L, f"{tag}_ini_get_string")(h, sb, kb, b"<dflt>", buf, 512)
        out[(s, k, "str")] = buf.value
    getattr(L, f"{tag}_ini_close")(h)
    return out


def test_facade_reader_matches_reference_reader(readers):
    libs, files = readers
    assert any(n.startswith("ref_") for n in files) or True
    for name, path in files.items():
        ref = _read_all(libs["ref"], "ref", path)
        fac = _read_all(libs["fac"], "fac", path)
        diff = {k: (ref[k], fac[k]) for k in ref if ref[k] != fac[k]}
        assert not diff, (name, diff)


def test_known_values(readers):
    libs, files = readers
    r = _read_all(libs["fac"], "fac", files["comments_and_quirks.ini"])
    assert r[("problem", "n0", "int")] == 16 and r[("problem", "n1", "int")] == 8
    assert r[("problem", "n2", "int")] == 12 and r[("problem", "n2", "str")] == b"12 # hash is not a comment here"
    assert r[("problem", "dt", "float")] == float(np.float32(0.001))
    assert r[("impl", "inplace", "bool_f")] == 0  # "true # trailing" is not a bool
    assert r[("optimization", "pref_wg_size", "str")] == b"continued value"
    assert r["error"] > 0


def test_adv_setup_from_reference_ini(readers):
    libs, files = readers
    if "ref_advection.ini" not in files:
        pytest.skip("reference advection.ini absent")
    fac = ctypes.CDLL(libs["fac"]._name)
    p = B.AdvParams()
    impl = ctypes.create_string_buffer(64)
    fac.fac_adv_setup.argtypes = [ctypes.c_char_p, ctypes.POINTER(B.AdvParams), ctypes.c_char_p, ctypes.c_int]
    err = fac.fac_adv_setup(files["ref_advection.ini"].encode(), ctypes.byref(p), impl, 64)
    assert err == 0
    assert (p.n0, p.n1, p.n2, p.maxIter) == (1024, 1024, 1024, 50)
    assert p.dt == float(np.float32(0.001)) and p.minRealVx == -1.0 and p.maxRealVx == 1.0
    assert p.pref_wg_size == 512 and p.gpu == 1 and impl.value == b"AdaptiveWg"
    assert p.dx == 1 / 1024 and p.dvx == 2 / 1024 and p.inv_dx == 1024


CONV_FIELDS = ("length", "channel_in", "channel_out", "k", "total_batch_size", "batch_size_n2", "n0", "n1", "n2",
               "n_write", "gpu", "pref_wg_size", "seq_size0", "seq_size2", "inplace")


def _conv_setup(libs, path):
    fac = ctypes.CDLL(libs["fac"]._name)
    fac.fac_conv_setup.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_ulonglong), ctypes.c_char_p,
                                   ctypes.c_int]
    out = (ctypes.c_ulonglong * 15)()
    impl = ctypes.create_string_buffer(64)
    err = fac.fac_conv_setup(path.encode(), out, impl, 64)
    return err, dict(zip(CONV_FIELDS, list(out))), impl.value


def _ref_reads(libs, path):
    """The raw values Conv1dParamsNonCopyable::setup (src/config/Conv1dParams.cpp:44-68) reads, through the
    reference's own reader, with the reference's defaults."""
    L = libs["ref"]
    h = L.ref_ini_open(path.encode())
    gi = lambda s, k, d: L.ref_ini_get_integer(h, s.encode(), k.encode(), d)  # noqa: E731
    gb = lambda s, k, d: L.ref_ini_get_bool(h, s.encode(), k.encode(), d)  # noqa: E731
    buf = ctypes.create_string_buffer(64)
    L.ref_ini_get_string(h, b"impl", b"kernelImpl", b"AdaptiveWg", buf, 64)
    r = {"length": gi("problem", "length", 1024), "channel_in": gi("problem", "channel_in", 3),
         "k": gi("problem", "k", 3), "total_batch_size": gi("problem", "total_batch_size", 262144),
         "batch_size_n2": gi("problem", "batch_size_n2", 512), "inplace": gb("impl", "inplace", 1),
         "gpu": gb("optimization", "gpu", 1), "pref_wg_size": gi("optimization", "pref_wg_size", 512),
         "seq_size0": gi("optimization", "seq_size0", 1), "seq_size2": gi("optimization", "seq_size2", 1),
         "kernelImpl": buf.value}
    L.ref_ini_close(h)
    return r


def test_conv_setup_from_reference_ini(readers):
    # the reference's conv1d.ini: every derived size is a known answer of Conv1dParams.cpp:44-68
    libs, files = readers
    if "ref_conv1d.ini" not in files:
        pytest.skip("reference conv1d.ini absent")
    err, p, impl = _conv_setup(libs, files["ref_conv1d.ini"])
    assert err == 0 and impl == b"AdaptiveWg"
    assert p == {"length": 1024, "channel_in": 3, "channel_out": 3, "k": 3, "total_batch_size": 262144,
                 "batch_size_n2": 512, "n0": 512, "n1": 3072, "n2": 512, "n_write": 1022, "gpu": 1,
                 "pref_wg_size": 512, "seq_size0": 1, "seq_size2": 1, "inplace": 1}
    ref = _ref_reads(libs, files["ref_conv1d.ini"])
    for k in ("length", "channel_in", "k", "total_batch_size", "batch_size_n2", "pref_wg_size", "seq_size0",
              "seq_size2", "gpu", "inplace"):
        assert p[k] == ref[k], k


def test_conv_setup_short_fields_wrap_like_the_reference(readers, tmp_path):
    # channel_in, k and pref_wg_size are `short unsigned` in Conv1dParams (Conv1dParams.hpp:16-31), so values
    # read as long wrap modulo 2^16: 65539 -> 3, 65541 -> 5, 66048 -> 512.
    libs, _ = readers
    f = tmp_path / "wrap.ini"
    f.write_text("[problem]\nlength = 100\nchannel_in = 65539\nk = 65541\ntotal_batch_size = 96\n"
                 "batch_size_n2 = 8\n[impl]\nkernelImpl = NDRange\ninplace = false\n"
                 "[optimization]\ngpu = no\npref_wg_size = 66048\nseq_size0 = 2\nseq_size2 = 4\n")
    err, p, impl = _conv_setup(libs, str(f))
    ref = _ref_reads(libs, str(f))
    assert err == 0 and impl == b"NDRange" == ref["kernelImpl"]
    assert (ref["channel_in"], ref["k"], ref["pref_wg_size"]) == (65539, 65541, 66048)
    assert (p["channel_in"], p["channel_out"], p["k"], p["pref_wg_size"]) == (3, 3, 5, 512)
    assert (p["n0"], p["n1"], p["n2"], p["n_write"]) == (12, 300, 8, 96)
    assert (p["gpu"], p["inplace"], p["seq_size0"], p["seq_size2"]) == (0, 0, 2, 4)


def test_conv_setup_defaults(readers, tmp_path):
    libs, _ = readers
    f = tmp_path / "empty.ini"
    f.write_text("[problem]\n")
    err, p, impl = _conv_setup(libs, str(f))
    assert err == 0 and impl == b"AdaptiveWg"
    assert p == {"length": 1024, "channel_in": 3, "channel_out": 3, "k": 3, "total_batch_size": 262144,
                 "batch_size_n2": 512, "n0": 512, "n1": 3072, "n2": 512, "n_write": 1022, "gpu": 1,
                 "pref_wg_size": 512, "seq_size0": 1, "seq_size2": 1, "inplace": 1}
