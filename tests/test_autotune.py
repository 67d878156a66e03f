"""Autotuner host logic (no GPU): workload signatures, the search space,
the user/shipped table lookup order and option application."""

import json

from fbme_amd import _lib, autotune


def test_plan_signature_is_a_multiset():
    a = autotune.plan_signature([(64, 64, 30, 40), (256, 256, 160, 160), (64, 64, 30, 40)], False)
    b = autotune.plan_signature([(64, 64, 30, 40), (64, 64, 30, 40), (256, 256, 160, 160)], False)
    assert a == b == "fast|64x64x30x40*2,256x256x160x160*1"
    assert autotune.plan_signature([(64, 64, 30, 40)], True).startswith("stable|")


def test_candidates_skip_tilings_without_lds_passes():
    only256 = autotune.plan_candidates([(256, 256, 160, 160)])
    # register passes everywhere: LDS tilings only vary under fft_kernel 1
    assert sum(c["fft_kernel"] == 0 for c in only256) == 1
    assert sum(c["fft_kernel"] == 1 for c in only256) == len(autotune.ROW_LINES) * len(autotune.COL_LINES)
    mixed = autotune.plan_candidates([(256, 256, 160, 160), (64, 128, 40, 90)])
    assert len(mixed) == 2 * len(autotune.ROW_LINES) * len(autotune.COL_LINES)


def test_lookup_prefers_user_table(tmp_path, monkeypatch):
    user = tmp_path / "user.json"
    shipped = tmp_path / "shipped.json"
    sig = autotune.plan_signature([(128, 128, 64, 64)], False)
    shipped.write_text(json.dumps({"gfx950": {sig: {"fft_kernel": 1, "row_lines": 2, "col_lines": 4}}}))
    monkeypatch.setattr(autotune, "SHIPPED", str(shipped))
    monkeypatch.setenv("FBME_TUNING", str(user))
    assert autotune.lookup("gfx950", sig)["row_lines"] == 2
    autotune.record("gfx950", sig, {"fft_kernel": 1, "row_lines": 8, "col_lines": 16})
    assert autotune.lookup("gfx950", sig)["row_lines"] == 8
    assert autotune.lookup("gfx942", sig) is None
    opt = autotune.apply_plan_entry(_lib.default_options(False), autotune.lookup("gfx950", sig))
    assert (opt.fft_kernel, opt.row_lines, opt.col_lines) == (1, 8, 16)


def test_shipped_table_is_well_formed():
    table = json.load(open(autotune.SHIPPED))
    for arch, entries in table.items():
        for sig, entry in entries.items():
            if sig.startswith("cg|"):
                assert entry["row_tile"] >= 0 and entry["col_tile"] >= 0
                assert entry["ms"] <= entry["default_ms"] * 1.0001
            else:
                assert sig.split("|")[0] in ("fast", "stable")
                assert entry["fft_kernel"] in (0, 1)
                assert entry["row_lines"] >= 1 and entry["col_lines"] >= 1
                assert entry["ms"] <= entry["default_ms"] * 1.0001
