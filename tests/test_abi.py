"""CPU checks of the C-ABI library: it loads, and it exports every function that
include/*.h declares (no GPU calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for fn in sorted(os.listdir(INCLUDE)):
        if not fn.endswith(".h"):
            continue
        src = open(os.path.join(INCLUDE, fn)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"//[^\n]*", "", src)
        src = re.sub(r"^\s*#[^\n]*", "", src, flags=re.M)
        src = src.replace('extern "C" {', "")
        # drop static inline bodies and typedefs
        src = re.sub(r"static\s+inline[^{]*\{.*?\n\}", "", src, flags=re.S)
        for stmt in src.split(";"):
            stmt = " ".join(stmt.split())
            if not stmt or stmt.startswith(("typedef", "#", "struct", "enum", "static")) or "(" not in stmt:
                continue
            m = re.match(r"^(?:extern \"C\" \{ )?[\w\s\*]+?\b(\w+)\s*\(", stmt)
            if m and not stmt.startswith("typedef"):
                names.add(m.group(1))
    return names


def test_headers_declare_the_reference_entry_points():
    names = declared_functions()
    for n in ("matchgate_brickwall_unitary_target_gradient_vector_hessian_matrix", "truncated_cg_hvp",
              "riemannian_trust_region_optimize_hvp", "optimize_matchgate_brickwall_circuit_hmat",
              "apply_matchgate_backward_array", "retract_matchgate", "mg_riemannian_gradient_hvp"):
        assert n in names


def test_library_exports_every_declared_symbol():
    import rqcopt_mg

    lib = ctypes.CDLL(rqcopt_mg.LIB_PATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_no_oracle_in_product_sources():
    """The product library neither links nor references the CPU oracle."""
    csrc = os.path.join(ROOT, "rqcopt_hpc-matchgate_amd", "csrc")
    for fn in os.listdir(csrc):
        txt = open(os.path.join(csrc, fn)).read()
        assert "orc_" not in txt and "mg_oracle" not in txt, fn
    py = open(os.path.join(ROOT, "rqcopt_hpc-matchgate_amd", "rqcopt_mg", "__init__.py")).read()
    assert "import oracle" not in py
