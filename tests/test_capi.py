"""The C ABI without a GPU: the library loads, exports every function that
include/dmgx.h declares, reports malformed nep.txt files as DMGX_EINVAL
before touching the device, and its host-side ensemble scalars reproduce the
reference's bitwise thermostat fixtures (tests/thermostats/fixtures)."""
import ctypes
import json
import os
import struct

import numpy as np
import pytest

from conftest import GOLDEN

FIX = os.path.join(GOLDEN, "thermostats")


def test_library_exports_every_declared_symbol(dmgx):
    L = dmgx._lib
    names = L.header_functions()
    assert len(names) >= 25
    so = ctypes.CDLL(L.LIB_PATH)
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing
    assert set(L._SIGNATURES) >= set(names)


BAD = {
    "unknown keyword": ("nep3 1 C\n", "not a supported model"),
    "symbol count": ("nep4 2 C\n", "symbols"),
    "cutoff tokens": ("nep4 1 C\ncutoff 4 3 50\n", "cutoff"),
    "mn too large": ("nep4 1 C\ncutoff 4 3 900 50\n", "819"),
    "l_max flags": ("nep4 1 C\ncutoff 4 3 50 40\nn_max 4 4\nbasis_size 8 8\nl_max 4 2 1 1\n", "not supported"),
    "missing params": ("nep4 1 C\ncutoff 4 3 50 40\nn_max 1 1\nbasis_size 1 1\nl_max 1 0 0\nANN 2 0\n0.1\n",
                       "missing"),
    "angular above radial": ("nep4 1 C\ncutoff 3 4 50 40\n", "angular cutoff"),
}


@pytest.mark.parametrize("case", sorted(BAD))
def test_malformed_nep_txt_is_einval(dmgx, tmp_path, case):
    text, msg = BAD[case]
    p = tmp_path / "nep.txt"
    p.write_text(text)
    h = ctypes.c_void_p()
    rc = dmgx.lib().dmgx_nep_create(str(p).encode(), ctypes.byref(h))
    assert rc == -1
    assert msg in dmgx.lib().dmgx_last_error().decode()
    assert not h.value


def test_missing_file_is_einval(dmgx, tmp_path):
    h = ctypes.c_void_p()
    assert dmgx.lib().dmgx_nep_create(str(tmp_path / "none.txt").encode(), ctypes.byref(h)) == -1


def _bits(x: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def test_svr_fixtures_bitwise(dmgx):
    L = dmgx.lib()
    cases = json.load(open(os.path.join(FIX, "svr.json")))["cases"]
    assert len(cases) == 16
    for c in cases:
        r = ctypes.c_void_p()
        assert L.dmgx_rng_create(c["seed"], ctypes.byref(r)) == 0
        got = [_bits(L.dmgx_gasdev(r)) for _ in range(c["gasdev_count"])]
        assert got == c["gasdev_bits"], c["seed"]
        for s in c["sumnoises"]:
            assert _bits(L.dmgx_resamplekin_sumnoises(s["nn"], r)) == s["bits"]
        for b in c["resamplekin_bits"]:
            assert _bits(L.dmgx_resamplekin(c["kk"], c["sigma"], c["ndeg"], c["taut"], r)) == b
        L.dmgx_rng_destroy(r)


def test_gamdev_rejects_shape_below_one(dmgx):
    L = dmgx.lib()
    r = ctypes.c_void_p()
    L.dmgx_rng_create(0, ctypes.byref(r))
    assert np.isnan(L.dmgx_gamdev(0, r))
    assert "ia >= 1" in L.dmgx_last_error().decode()
    L.dmgx_rng_destroy(r)


def test_nhc_fixtures_bitwise(dmgx):
    L = dmgx.lib()
    hx = lambda s: struct.unpack(">d", bytes.fromhex(s))[0]
    cases = json.load(open(os.path.join(FIX, "nhc.json")))["cases"]
    ramps = {"case0": [300.0] * 4, "case1": [725.0, 700.0, 675.0, 650.0],
             "case2": [1200.0, 1225.0, 1250.0, 1275.0]}
    for c in cases:
        st = dmgx._lib.Nhc()
        assert L.dmgx_nhc_init(c["N"], c["T0"], c["Tdamp"], c["dt"], ctypes.byref(st)) == 0
        assert [_bits(v) for v in st.mass] == [_bits(hx(v)) for v in c["initial_mass"]]
        ek2 = c["initial_Ek2"]
        for k, s in enumerate(c["steps"]):
            assert s["target_temperature"] == ramps[c["name"]][k]
            f = L.dmgx_nhc_half_step(ctypes.byref(st), ek2, 8.617343e-5 * s["target_temperature"],
                                     3.0 * c["N"], 0.5 * c["dt"])
            ek2 *= f * f
            assert _bits(f) == _bits(hx(s["factor"]))
            assert _bits(ek2) == _bits(hx(s["Ek2_after"]))
            for key in ("position", "momentum", "mass"):
                assert [_bits(v) for v in getattr(st, key)] == [_bits(hx(v)) for v in s[key]], key


def test_npt_scr_scale_closed_form(dmgx):
    L = dmgx.lib()
    pt, pc, pi, T, V, g = 1e-4 / 160.2177, 0.01, 2e-4, 300.0, 1000.0, 0.37
    expect = (1.0 - pc * (pt - pi)) + np.sqrt(0.666666666666667 * pc * 8.617343e-5 * T / V) * g
    assert L.dmgx_npt_scr_scale(pt, pc, pi, T, V, g) == expect
