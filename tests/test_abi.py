"""The C ABI boundary without a GPU: libphotonprop_hip.so loads, exports every
function include/photonprop_hip.h declares, the ctypes POD layouts match the
header, and calls fail cleanly (status + message) rather than crash."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "photonprop_hip.h")
HEADERS = sorted(os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))
                 if f.endswith(".h"))


def declared_functions():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(pp_[a-z_0-9]+)\s*\(", text))
    return sorted(names)


def test_header_declares_the_boundary():
    fns = declared_functions()
    for name in ("pp_create", "pp_destroy", "pp_set_medium", "pp_set_spectrum",
                 "pp_set_targets", "pp_propagate", "pp_copy_hits", "pp_last_error"):
        assert name in fns


def test_library_exports_every_declared_symbol(P):
    lib = P._lib.load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", P._lib.library_path()],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (pp_[a-z_0-9]+)", out))
    assert set(declared_functions()) <= exported
    bound = {s[0] for s in P._lib.SIGNATURES}
    assert set(declared_functions()) == bound


def test_pod_layouts_match_header(P):
    """Compile a probe against the header and compare sizes/offsets."""
    probe = r"""
#include <stddef.h>
#include <stdio.h>
#include "photonprop_hip.h"
#include "photonprop_surrogate.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu\n", sizeof(pp_source), sizeof(pp_medium), sizeof(pp_target),
         sizeof(pp_config), sizeof(pp_cascade), sizeof(pp_sur_domain), sizeof(pp_adam));
  printf("%zu %zu %zu %zu\n", offsetof(pp_source, photons), offsetof(pp_source, long_a),
         offsetof(pp_medium, pressure), offsetof(pp_target, module_id));
  return 0;
}
"""
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(c, "w").write(probe)
        subprocess.run(["gcc", "-I", os.path.dirname(HEADER), "-o", exe, c], check=True)
        lines = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    sizes = [int(x) for x in lines[0].split()]
    offs = [int(x) for x in lines[1].split()]
    L = P._lib
    assert sizes == [ctypes.sizeof(L.pp_source), ctypes.sizeof(L.pp_medium),
                     ctypes.sizeof(L.pp_target), ctypes.sizeof(L.pp_config),
                     ctypes.sizeof(L.pp_cascade), ctypes.sizeof(L.pp_sur_domain),
                     ctypes.sizeof(L.pp_adam)]
    assert offs == [L.pp_source.photons.offset, L.pp_source.long_a.offset,
                    L.pp_medium.pressure.offset, L.pp_target.module_id.offset]


def test_record_bytes_and_version(P):
    lib = P._lib.load_library()
    assert lib.pp_abi_version() == 1
    # 64-byte records, the reference's sizeof(PhotonHit) (src/photon_prop_cuda.jl:872-875)
    assert lib.pp_hit_record_bytes() == 64


def test_create_without_device_fails_cleanly(P):
    lib = P._lib.load_library()
    if lib.pp_device_count() > 0:
        pytest.skip("a device is visible; covered by the GPU tests")
    h = ctypes.c_void_p()
    rc = lib.pp_create(ctypes.byref(h), 0, 1024, None)
    assert rc == P._lib.PP_ERR_INVALID
    assert "no HIP device" in P._lib.last_error()
    assert not h.value


def test_null_arguments_fail_cleanly(P):
    lib = P._lib.load_library()
    assert lib.pp_set_medium(None, None) == P._lib.PP_ERR_INVALID
    assert lib.pp_propagate(None, None, 0, 0, 15, 1.0, None, None) == P._lib.PP_ERR_INVALID
    assert lib.pp_copy_hits(None, 0, 0, *([None] * 5), 0, *([None] * 5)) == P._lib.PP_ERR_INVALID
    assert lib.pp_destroy(None) == 0


def test_product_has_no_oracle_dependency():
    """The shipped library and package never reference the oracle."""
    pkg = os.path.join(ROOT, "photonpropagation.jl_amd")
    for dp, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                text = open(os.path.join(dp, f)).read()
                assert "oracle" not in text.lower() or f == "__init__.py" and False, f
    out = subprocess.run(["ldd", os.path.join(pkg, "photonprop", "libphotonprop_hip.so")],
                         capture_output=True, text=True).stdout
    assert "pp_oracle" not in out


def test_host_spectrum_table(P):
    """make_cuda_spectral_dist of a flat spectrum is linear in u; the Cherenkov
    table is monotonic within the range and matches the analytic inverse CDF
    of 1/wl^2 weighting to the knot resolution."""
    flat = P.make_spectral_dist(lambda w: np.ones_like(w), (300.0, 700.0))
    tab = P.make_cuda_spectral_dist(flat, (300.0, 700.0), 600).table
    np.testing.assert_allclose(tab, np.linspace(300, 700, 600), atol=1e-3)
    inv2 = P.make_spectral_dist(lambda w: 1.0 / w**2, (300.0, 600.0))
    tab = P.make_cuda_spectral_dist(inv2, (300.0, 600.0), 600).table
    u = np.linspace(0, 1, 600)
    exact = 1.0 / (1 / 300.0 - u * (1 / 300.0 - 1 / 600.0))  # inverse CDF of 1/wl^2
    np.testing.assert_allclose(tab, exact, atol=0.05)
    med = P.CascadiaMediumProperties()
    spec = P.make_cherenkov_spectrum((300.0, 800.0), med)
    t = P.make_cuda_spectral_dist(spec.spectral_dist, (300.0, 800.0), 600).table
    assert t[0] == 300.0 and abs(t[-1] - 800.0) < 1e-3 and np.all(np.diff(t) > 0)
    # Frank-Tamm yield per metre of track in 300-800 nm: ~43e3 photons/m
    assert 4.0e4 < spec.spectral_dist.normalization < 4.6e4
