"""The tile-shape autotuner's host logic (no GPU): candidate bookkeeping and
the rewrite of tile_config.hpp."""
from fv3core_amd import autotune


def test_config_has_a_default_for_every_tuned_macro():
    text = autotune.CONFIG.read_text()
    for _, _, macros, shapes in autotune.KERNELS.values():
        for macro in macros:
            assert f"#ifndef {macro}\n#define {macro} " in text
        assert len(set(shapes)) == len(shapes)


def test_best_shape_is_the_fastest_and_rewrite_touches_only_its_macros():
    results = {"delnflux": {"8x4": 84.5, "16x4": 97.0, "4x4": 108.9},
               "a2b": {"16x8": 67.4, "16x4": 75.3}}
    best = autotune.best_shapes(results)
    assert best == {"delnflux": (8, 4), "a2b": (16, 8)}
    text = autotune.CONFIG.read_text()
    swapped = autotune.rewrite_config(text, {"delnflux": (12, 2)})
    assert "#define FV3_DN_TJ 12" in swapped and "#define FV3_DN_NY 2" in swapped
    # the other kernel's defaults and the rest of the file are untouched
    assert swapped.replace("FV3_DN_TJ 12", "").replace("FV3_DN_NY 2", "") == \
        text.replace(f"FV3_DN_TJ {_default(text, 'FV3_DN_TJ')}", "").replace(
            f"FV3_DN_NY {_default(text, 'FV3_DN_NY')}", "")
    # rewriting with the current defaults is the identity
    cur = {k: (_default(text, m[0]), _default(text, m[1]))
           for k, (_, _, m, _) in autotune.KERNELS.items()}
    assert autotune.rewrite_config(text, cur) == text


def _default(text, macro):
    import re
    return int(re.search(rf"#define {macro} (\d+)", text).group(1))
