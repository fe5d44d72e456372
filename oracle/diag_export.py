"""DiagnosticFilm queries and exports, restated over a film array.

TEST INFRASTRUCTURE ONLY: imported by tests/.  The product never imports it.

The film is a (sampled_h, sampled_w) array of oracle.PIXEL_DT pixels
(PixelPathData<16, 4>), as the oracle's diagnostic render and the product's
get_diagnostic_film_bytes() return it.  Each function restates one query of
the reference (LucidRenderer/cpp_renderer/include/diagnostics/):

* global_stats        DiagnosticFilm::global_stats        diagnostic_film.hpp:135-155
* top_groups          top_groups_by_variance / _by_mean   diagnostic_film.hpp:160-224
* group_info          DiagnosticExporter::convert_refs    diagnostic_export.hpp:254-284
* coarse_stats        DiagnosticExporter::get_coarse_stats diagnostic_export.hpp:109-155
* metadata_json       DiagnosticFilm::metadata_json       diagnostic_film.hpp:230-244
* export_binary       DiagnosticFilm::export_binary "PVAR" diagnostic_film.hpp:247-299
* suggestions         DiagnosticExporter::generate_suggestions diagnostic_export.hpp:188-249
* variance_map        Renderer::get_diagnostic_variance_map pybind_renderer.hpp:941-960

Float arithmetic is numpy float32 in the reference's operation order
(PathStatistics::total_variance / mean_luminance / variance_luminance,
path_stats.hpp:103-123), so results are comparable bit for bit with a
float32 implementation that does not fuse multiply-adds.

The reference sorts with std::partial_sort, which leaves the order of equal
keys unspecified; top_groups returns the groups in descending key order with
ties kept in film scan order, and `ties` (the key of the n-th entry) lets a
test accept any order among equal keys.

How it is pinned (tests/test_oracle_diag_export.py): the reference's own
export tests (test_diagnostic_export.cpp: PVAR magic, JSON fields, glossy
top group, 8 coarse types, suggestions for high-variance glossy groups),
the PVAR length 28 + 22 per sampled pixel and its fields decoded with
struct on hand-built films, brute-force sorting for the top groups, and the
coarse-stat merge against the count-weighted mean computed in float64.
"""

from __future__ import annotations

import numpy as np

import oracle

F = np.float32
LUM = (F(0.2126), F(0.7152), F(0.0722))

BSDF_SHORT = "DGMTLE"     # path_types.hpp:56-66
BSDF_HECKBERT = "DSSSEE"  # :72-82
LIGHT_HECKBERT = {1: "LSD", 2: "LDD", 3: "LSD", 4: "LSD", 5: "LDE", 0: "L??"}  # :147-157
COARSE_NAMES = ["Direct-NoDelta-BSDF", "Direct-NoDelta-NEE", "Direct-Delta-BSDF", "Direct-Delta-NEE",
                "Indirect-NoDelta-BSDF", "Indirect-NoDelta-NEE", "Indirect-Delta-BSDF",
                "Indirect-Delta-NEE"]  # path_stats.hpp:26-41
STRATEGY_NAMES = ["BSDF", "NEE", "MIS_BSDF", "MIS_NEE"]  # path_types.hpp:113-121


def total_variance(var) -> np.float32:
    return (F(var[0]) + F(var[1])) + F(var[2])


def mean_luminance(mean) -> np.float32:
    return (LUM[0] * F(mean[0]) + LUM[1] * F(mean[1])) + LUM[2] * F(mean[2])


def variance_luminance(var) -> np.float32:
    return ((LUM[0] * LUM[0]) * F(var[0]) + (LUM[1] * LUM[1]) * F(var[1])) + (LUM[2] * LUM[2]) * F(var[2])


def signature(g) -> str:
    return "".join(BSDF_SHORT[t] if t < 6 else "?" for t in g["vertices"]["bsdf_type"][: int(g["depth"])])


def signature_heckbert(g) -> str:
    types = [int(t) for t in g["vertices"]["bsdf_type"][: int(g["depth"])]]
    s = LIGHT_HECKBERT.get(int(g["light_type"]), "L??")
    end = len(types)
    while end > 0 and types[end - 1] in (4, 5):
        end -= 1
    if end > 0:
        s += " " + " ".join(BSDF_HECKBERT[t] if t < 6 else "?" for t in types[:end])
    return s + " DSE"


def global_stats(film) -> dict:
    px = film.reshape(-1)
    act = px["total_samples"] > 0
    return {"total_samples": int(px["total_samples"][act].sum()), "active_pixels": int(act.sum()),
            "total_groups": int(px["num_groups"][act].sum()), "total_overflow": int(px["overflow_count"][act].sum())}


def top_groups(film, n: int, subsample: int, by_mean: bool = False):
    """[(pixel_x, pixel_y, group record)] of the n best groups, and the key of
    the last one kept (None if none)."""
    sh, sw = film.shape
    min_count = 1 if by_mean else 2
    cand = []
    for sy in range(sh):
        for sx in range(sw):
            p = film[sy, sx]
            for k in range(int(min(p["num_groups"], 16))):
                g = p["groups"][k]
                if int(g["stats"]["count"]) >= min_count:
                    key = mean_luminance(g["stats"]["mean"]) if by_mean else total_variance(g["stats"]["variance"])
                    cand.append((key, sx * subsample, sy * subsample, g))
    order = sorted(range(len(cand)), key=lambda i: -float(cand[i][0]))  # stable: ties in scan order
    keep = [cand[i] for i in order[:n]]
    return [(x, y, g) for _, x, y, g in keep], (keep[-1][0] if keep else None), cand


def group_info(x, y, g) -> dict:
    s = g["stats"]
    return {"pixel_x": int(x), "pixel_y": int(y), "signature": signature(g), "sample_count": int(s["count"]),
            "mean_luminance": mean_luminance(s["mean"]), "variance_luminance": variance_luminance(s["variance"]),
            "mean_rgb": [F(v) for v in s["mean"]], "variance_rgb": [F(v) for v in s["variance"]],
            "depth": int(g["depth"]), "coarse_type": int(g["coarse_type"]),
            "coarse_type_name": COARSE_NAMES[int(g["coarse_type"]) & 7]}


def coarse_stats(film, subsample: int) -> list:
    """The reference's merge, including its fixed pixel_data(2 sx, 2 sy)
    lookup whatever the film's subsample factor (diagnostic_export.hpp:118-121)."""
    sh, sw = film.shape
    cnt = [0] * 8
    mean = [[F(0), F(0), F(0)] for _ in range(8)]
    var = [[F(0), F(0), F(0)] for _ in range(8)]
    for sy in range(sh):
        for sx in range(sw):
            qx, qy = (2 * sx) // subsample, (2 * sy) // subsample  # pixel_data(x, y) -> sampled index
            if qx >= sw or qy >= sh:
                continue
            p = film[qy, qx]
            for t in range(8):
                cs = p["coarse"][t]
                c = int(cs["count"])
                if c == 0:
                    continue
                cnt[t] = (cnt[t] + c) & 0xFFFFFFFF
                w = F(F(c) / F(cnt[t]))
                for ch in range(3):
                    mean[t][ch] = F(mean[t][ch] * F(F(1) - w)) + F(F(cs["mean"][ch]) * w)
                    var[t][ch] = max(var[t][ch], F(cs["variance"][ch]))
    out = []
    for t in range(8):
        out.append({"type_id": t, "name": COARSE_NAMES[t], "sample_count": cnt[t],
                    "mean_luminance": mean_luminance(mean[t]), "variance_luminance": variance_luminance(var[t]),
                    "mean_rgb": list(mean[t]), "variance_rgb": list(var[t])})
    return out


def metadata_json(width, height, subsample, max_groups, max_outliers, total_recorded) -> str:
    sw, sh = width // subsample, height // subsample
    return ("{\n"
            f"  \"width\": {width},\n"
            f"  \"height\": {height},\n"
            f"  \"sampled_width\": {sw},\n"
            f"  \"sampled_height\": {sh},\n"
            f"  \"subsample_factor\": {subsample},\n"
            f"  \"max_groups\": {max_groups},\n"
            f"  \"max_outliers\": {max_outliers},\n"
            f"  \"total_samples\": {total_recorded}\n"
            "}\n")


PVAR_HEADER_DT = np.dtype([("magic", "S4"), ("version", "<u4"), ("width", "<u4"), ("height", "<u4"),
                           ("sampled_width", "<u4"), ("sampled_height", "<u4"), ("subsample", "u1"),
                           ("max_groups", "u1"), ("max_outliers", "u1"), ("pad", "u1")])
PVAR_PIXEL_DT = np.dtype([("group_count", "<u2"), ("total_samples", "<u4"), ("mean", "<f4", 3),
                          ("variance", "<f4")])  # packed: 22 bytes
assert PVAR_HEADER_DT.itemsize == 28 and PVAR_PIXEL_DT.itemsize == 22


def export_binary(film, width, height, subsample, max_groups, max_outliers) -> bytes:
    sh, sw = film.shape
    hdr = np.zeros(1, PVAR_HEADER_DT)
    hdr[0] = (b"PVAR", 1, width, height, sw, sh, subsample, max_groups, max_outliers, 0)
    px = film.reshape(-1)
    rec = np.zeros(px.shape[0], PVAR_PIXEL_DT)
    rec["group_count"] = (px["num_groups"] & 0xFFFF).astype(np.uint16)
    rec["total_samples"] = (px["total_samples"] & 0xFFFFFFFF).astype(np.uint32)
    rec["mean"] = px["total"]["mean"]
    v = px["total"]["variance"].astype(np.float32)
    rec["variance"] = (v[:, 0] + v[:, 1]) + v[:, 2]
    return hdr.tobytes() + rec.tobytes()


_SUGGEST = [
    ("Settings", "warning", "High variance detected in glossy reflections",
     "Consider increasing samples or using adaptive sampling for glossy materials"),
    ("Algorithm", "info", "Significant variance from indirect illumination",
     "MIS algorithm recommended; consider path guiding for complex scenes"),
    ("Scene", "info", "High variance in diffuse lighting",
     "Check for small or distant light sources; consider area lights"),
    ("Settings", "warning", "Many path types exceeded tracking capacity",
     "Consider using 'Detailed' preset for more accurate analysis"),
]


def suggestions(film, subsample: int) -> list:
    top, _, _ = top_groups(film, 10, subsample)
    if not top:
        return []
    glossy = indirect = diffuse = 0
    for _, _, g in top:
        types = [int(t) for t in g["vertices"]["bsdf_type"][: int(g["depth"])]]
        has_glossy, has_indirect = 1 in types, int(g["depth"]) > 2
        glossy += has_glossy
        indirect += has_indirect
        diffuse += (not has_glossy and not has_indirect)
    out = []
    if glossy > 3:
        out.append(_SUGGEST[0])
    if indirect > 3:
        out.append(_SUGGEST[1])
    if diffuse > 3:
        out.append(_SUGGEST[2])
    st = global_stats(film)
    if st["total_samples"] > 0 and st["total_overflow"] > st["active_pixels"] * 0.1:
        out.append(_SUGGEST[3])
    return out


def variance_map(film, width, height, subsample) -> np.ndarray:
    sh, sw = film.shape
    v = film["total"]["variance"].astype(np.float32)
    tv = (v[..., 0] + v[..., 1]) + v[..., 2]
    out = np.zeros((height, width), np.float32)
    ys, xs = np.arange(height) // subsample, np.arange(width) // subsample
    ok_y, ok_x = ys < sh, xs < sw
    out[np.ix_(ok_y, ok_x)] = tv[np.ix_(ys[ok_y], xs[ok_x])]
    return out.reshape(-1)


def film_from_bytes(raw, sampled_h, sampled_w):
    return np.ascontiguousarray(raw).view(np.uint8).reshape(-1)[: sampled_h * sampled_w * oracle.PIXEL_DT.itemsize] \
        .view(oracle.PIXEL_DT).reshape(sampled_h, sampled_w)
