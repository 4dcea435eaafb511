"""CPU checks of the C-ABI: the library loads and exports every symbol the headers declare."""
import os
import re

import electricbubbles_amd as eb
from tests.conftest import ROOT


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            txt = open(os.path.join(inc, f)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names.update(re.findall(r"\b(eb_[a-z0-9_]+)\s*\(", txt))
    return names


def test_headers_declare_what_python_binds():
    assert set(eb.EXPORTS) <= declared_symbols()


def test_library_exports_every_declared_symbol():
    lib = eb.lib()
    missing = [s for s in sorted(declared_symbols()) if not hasattr(lib, s)]
    assert not missing, missing


def test_nterms_rule():
    # NTERMS = -NINT(log(TOL)/log(2)) (cadapdrop.f:126)
    assert eb.nterms_for(1e-13) == 43
    assert eb.nterms_for(1e-10) == 33
    assert eb.nterms_for(1e-12) == 40
