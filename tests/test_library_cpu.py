"""CPU checks of the C-ABI library: it loads, exports every symbol that
include/channelflow.h declares, and rejects bad parameters before touching
a device.  No compute call is made here."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "channelflow.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cf_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_binding_surface():
    import channelflow as C
    assert set(declared_symbols()) == set(C.EXPORTS)


def test_library_exports_every_declared_symbol():
    import channelflow as C
    lib = C.load_library()
    out = subprocess.run(["nm", "-D", "--defined-only", C.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\sT\s(cf_[a-z0-9_]+)", out))
    for name in declared_symbols():
        assert name in exported, name
        assert getattr(lib, name) is not None


@pytest.mark.parametrize("kw", [dict(nnx=48, nny=48), dict(nnx=64, nny=32), dict(nnz=2),
                                dict(imodel=5), dict(isurface=7), dict(isurface=0), dict(dt=0.0),
                                dict(precision=7)])
def test_create_rejects_unsupported_parameters(kw):
    import channelflow as C
    args = dict(nnx=32, nny=32, nnz=16)
    args.update(kw)
    with pytest.raises(C.ChannelFlowError) as ei:
        C.ChannelFlow(C.RunParams(**args))
    assert ei.value.code == C.CF_E_ARG


def test_product_has_no_oracle_dependency():
    """The library and its binding never import, link or read the oracle."""
    pkg = os.path.join(ROOT, "paes-channelflow_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".h", ".F90", ".f90")):
                txt = open(os.path.join(dirpath, f), errors="ignore").read()
                assert "cf_oracle" not in txt and "oracle/" not in txt, f


def test_kernel_families_of_the_timing_api():
    """cf_time_steps reports CF_NKERN kernel families; the header's constant,
    the binding's and the library's names agree (host calls only)."""
    import channelflow as C
    lib = C.load_library()
    m = re.search(r"#define\s+CF_NKERN\s+(\d+)", open(HEADER).read())
    assert m and int(m.group(1)) == C.CF_NKERN
    names = [lib.cf_kernel_name(k).decode() for k in range(C.CF_NKERN)]
    assert names == ["k_wall", "k_col_synth", "k_divz_band", "k_rows", "k_col_anal", "k_sgs_cols_in",
                     "k_sgs_rows", "k_sgs_cols_out", "k_press", "k_erows", "k_reduce_part"]
    assert lib.cf_kernel_name(C.CF_NKERN).decode() == "" and lib.cf_kernel_name(-1).decode() == ""
