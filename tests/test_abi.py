"""The C-ABI library loads on the CPU and exports every symbol of include/refrakt_amd.h."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "refrakt_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(rf_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(rf):
    L = rf.lib()
    declared = _declared()
    assert len(declared) > 40
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    # the ctypes binding covers the whole header
    from refrakt_amd import api
    assert sorted(api.exported_symbols()) == declared


def test_struct_layouts_match(rf):
    from refrakt_amd import api
    L = rf.lib()
    for i, st in enumerate(api._STRUCTS):
        assert L.rf_struct_size(i) == api.C.sizeof(st)


def test_variation_registry_matches_oracle_table(rf, orc):
    names = rf.variation_names()
    assert len(names) == 76
    L = orc.lib()
    onames = [L.orc_var_name(i).decode() for i in range(L.orc_var_count())]
    assert sorted(names) == sorted(onames)
    for n in names:
        assert rf.variation_params(n) == orc.var_params(n), n
    assert rf.variation_id("no_such_variation") == -1
    # flags: radial_blur skips the weight multiply, pre_blur is applied before the precalcs
    L2 = rf.lib()
    assert L2.rf_variation_flags(rf.variation_id("radial_blur")) & 1
    assert L2.rf_variation_flags(rf.variation_id("pre_blur")) & 2
    assert L2.rf_variation_flags(rf.variation_id("julia")) & 0x100
    assert not L2.rf_variation_flags(rf.variation_id("linear")) & 0x100


def test_generated_tables_are_current():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "refrakt_amd", "gen_variations.py"), "--check"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


def test_context_creation_fails_loudly_without_gpu(rf):
    import torch
    if torch.cuda.is_available():
        return
    try:
        rf.Context(1 << 10, 1)
    except rf.RefraktError as e:
        assert "rf_create" in str(e)
    else:
        raise AssertionError("rf_create succeeded without a GPU")


def test_create_rejects_bad_sim_parameters(rf):
    """Argument checks run before any HIP call, so they hold on the CPU too."""
    from refrakt_amd import api
    L = rf.lib()
    for kw, word in [({"block_waves": 16}, "block_waves"), ({"engine": 7}, "engine"),
                     ({"hist_mode": 9}, "hist mode")]:
        args = dict(total_particles=1 << 16, temporal_samples=1, shuffle_count=0, seed=1, device=0,
                    engine=api.RF_ENGINE_FAST, hist_mode=1, block_waves=4)
        args.update(kw)
        sp = api.SimParams(**args)
        out = api.C.c_void_p()
        assert L.rf_create(api.C.byref(sp), api.C.byref(out)) == -1  # RF_ERR_ARG
        assert word in L.rf_last_error(None).decode()
        assert not out.value


def test_division_rewrite_keeps_c_grouping():
    """Variation sources divide with `/`; the generator turns every a / b into
    a * rf_rcp(b).  `/` and `*` share precedence and associativity, so only
    the divisor's extent matters: a unary expression (casts, prefix signs,
    calls, subscripts, parenthesised groups), nested divisions included."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("gen_variations", os.path.join(ROOT, "refrakt_amd",
                                                                                 "gen_variations.py"))
    g = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(g)
    cases = {
        "ox = x / d; oy = y / d;": "ox = x * rf_rcp(d); oy = y * rf_rcp(d);",
        "a * sinf(t) / PI": "a * sinf(t) * rf_rcp(PI)",
        "k = 2.0f / (r + 1.0f);": "k = 2.0f * rf_rcp((r + 1.0f));",
        "m / -b[2] + c": "m * rf_rcp(-b[2]) + c",
        "p / (float)k / q": "p * rf_rcp((float)k) * rf_rcp(q)",
        "e / fmaxf(u * u, 1e-20f)": "e * rf_rcp(fmaxf(u * u, 1e-20f))",
        "c / (1.0f + c * x / r) / r": "c * rf_rcp((1.0f + c * x * rf_rcp(r))) * rf_rcp(r)",
        "(int)((rr + dx) / (2.0f * dx))": "(int)((rr + dx) * rf_rcp((2.0f * dx)))",
        "y / 2.0f": "y * rf_rcp(2.0f)",
        "v / 1.5e-3f;": "v * rf_rcp(1.5e-3f);",
    }
    for src, want in cases.items():
        assert g.rcp_divisions(src) == want, (src, g.rcp_divisions(src))
    with pytest.raises(ValueError):
        g.rcp_divisions("x /= d;")
