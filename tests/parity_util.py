"""Agreement statistics between a GPU render and the oracle render of the same scene."""
import numpy as np


def compare(gpu, ref):
    m_g = gpu["mask"].astype(bool)
    m_r = ref["mask"].astype(bool)
    both = m_g & m_r
    n = m_g.size
    out = {"pixels": n, "hits_ref": int(m_r.sum()), "mask_agree": float((m_g == m_r).mean())}
    out["iters_agree"] = float((gpu["iters"] == ref["iters"]).mean())
    out["iters_agree_hits"] = float((gpu["iters"][both] == ref["iters"][both]).mean()) if both.any() else 1.0
    if both.any():
        d_g = gpu["depth"][both].astype(np.float64)
        d_r = ref["depth"][both].astype(np.float64)
        rel = np.abs(d_g - d_r) / np.abs(d_r)
        out["depth_rel_median"] = float(np.median(rel))
        out["depth_rel_p999"] = float(np.quantile(rel, 0.999))
        out["depth_within_1e-3"] = float((rel <= 1e-3).mean())
        out["depth_within_1e-4"] = float((rel <= 1e-4).mean())
    err = np.abs(gpu["rgb"].astype(np.float64) - ref["rgb"].astype(np.float64)).max(axis=-1)
    out["rgb_abs_median"] = float(np.median(err))
    out["rgb_within_1e-2"] = float((err <= 1e-2).mean())
    out["rgb_within_1e-3"] = float((err <= 1e-3).mean())
    if "colour" in gpu and "colour" in ref:
        dc = np.abs(gpu["colour"].astype(np.int32) - ref["colour"].astype(np.int32)).max(axis=-1)
        out["colour_equal"] = float((dc == 0).mean())
        out["colour_within_1"] = float((dc <= 1).mean())
    if "opacity" in gpu and "opacity" in ref:
        do = np.abs(gpu["opacity"].astype(np.int64) - ref["opacity"].astype(np.int64))
        out["opacity_max_diff"] = int(do.max())
        out["opacity_within_2"] = float((do <= 2).mean())
    if "post_rgb" in gpu and "post_rgb" in ref:
        e = np.abs(gpu["post_rgb"].astype(np.float64) - ref["post_rgb"].astype(np.float64)).max(axis=-1)
        out["post_within_1e-2"] = float((e <= 1e-2).mean())
        out["post_within_1e-3"] = float((e <= 1e-3).mean())
    return out
