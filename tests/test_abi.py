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
