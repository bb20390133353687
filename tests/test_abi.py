"""C ABI of libpincflow_hip.so without a GPU: the library loads, exports every
symbol include/pincflow_hip.h declares, and the ctypes mirror of pcf_config
matches the C layout and defaults."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "pincflow_hip.h").read_text()
    return sorted(set(re.findall(r"\b(pcf_[a-z_0-9]+)\s*\(", text)))


def test_header_symbols_are_exported(abi):
    lib = abi.load_library()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_signatures_cover_header(abi):
    covered = {name for name, _, _ in abi.SIGNATURES}
    assert set(declared_symbols()) <= covered


def test_config_layout_and_defaults(pcf, abi):
    lib = abi.load_library()
    assert lib.pcf_config_size() == ctypes.sizeof(abi.PcfConfig)
    c = abi.PcfConfig()
    lib.pcf_config_default(ctypes.byref(c))
    mine = pcf.to_config(pcf.Namelists())
    for name, _ in abi.PcfConfig._fields_:
        if name in ("hills",):
            continue
        assert getattr(c, name) == getattr(mine, name), name


def test_oracle_config_layout(abi, oracle_mod):
    assert oracle_mod.lib().orc_config_size() == ctypes.sizeof(abi.PcfConfig)


def test_invalid_options_raise(pcf):
    nl = pcf.Namelists()
    nl.atmosphere.model = "Anelastic"
    with pytest.raises(ValueError):
        pcf.to_config(nl)
    nl = pcf.Namelists()
    nl.atmosphere.background = "Bogus"
    with pytest.raises(ValueError):
        pcf.to_config(nl)


def _header_config_fields():
    """(name, Julia type) of every pcf_config member, in order, from the header."""
    import re
    src = (ROOT / "include" / "pincflow_hip.h").read_text()
    body = re.search(r"typedef struct pcf_config \{(.*?)\} pcf_config;", src, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    out = []
    for stmt in body.split(";"):
        stmt = stmt.strip()
        if not stmt:
            continue
        m = re.match(r"(int32_t|double|pcf_hill)\s+(.*)", stmt, re.S)
        assert m, stmt
        for d in m.group(2).split(","):
            d = d.strip()
            arr = re.match(r"(\w+)\[PCF_MAX_HILLS\]", d)
            if arr:
                out.append((arr.group(1), "NTuple{32,PcfHill}"))
            else:
                out.append((d, {"int32_t": "Int32", "double": "Float64"}[m.group(1)]))
    return out


def test_julia_config_mirror_matches_the_header(abi):
    """The Julia stub's PcfConfig in INTEGRATION.md has the header's fields in
    the header's order with the matching types, and the ctypes mirror too."""
    import re
    text = (ROOT / "INTEGRATION.md").read_text()
    block = re.search(r"\nstruct PcfConfig\n(.*?)\nend\n", text, re.S).group(1)
    julia = [tuple(x.strip() for x in line.split("::")) for line in block.split("\n")
             if line.strip() and not line.strip().startswith("#")]
    header = _header_config_fields()
    assert julia == header
    assert [n for n, _ in abi.PcfConfig._fields_] == [n for n, _ in header]


def test_network_sizes_are_checked_before_any_device_call(abi):
    """pcf_net_create refuses grids it cannot index (nx not a multiple of 32,
    more than 2^31 - 1 voxels) with PCF_E_ARG and a message, before it
    touches a device"""
    lib = abi.load_library()
    h = ctypes.c_void_p()
    for nx, ny, nz, words in ((48, 8, 8, b"multiple of 32"), (2048, 2048, 1024, b"2^31")):
        assert lib.pcf_net_create(nx, ny, nz, 0, ctypes.byref(h)) == abi.E_ARG
        assert words in lib.pcf_last_error()
