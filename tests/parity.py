"""Helpers shared by the GPU parity tests: run a variant through the C ABI and its FP64
oracle on the same input, and summarise the difference."""
import json
import os
import sys

import numpy as np

# JSON-lines file the parity records are appended to (unset: printed only)
REPORT = os.environ.get("MPX_PARITY_REPORT")


def psnr(err):
    mse = float(np.mean(err ** 2))
    return float("inf") if mse == 0 else 10 * np.log10(1.0 / mse)


def report(case, **fields):
    """Print one parity record, and append it to $MPX_PARITY_REPORT when that is set."""
    rec = dict(case=case, **fields)
    line = json.dumps(rec, sort_keys=True, default=float)
    print("PARITY " + line, flush=True)
    if REPORT:
        try:
            os.makedirs(os.path.dirname(os.path.abspath(REPORT)), exist_ok=True)
            with open(REPORT, "a") as f:
                f.write(line + "\n")
        except OSError:
            pass
    sys.stdout.flush()


def compare(case, got, ref, got_rows=None, ref_rows=None, site_shape=None, site_to_out=None):
    """Max-abs / PSNR over all pixels, bucket agreement, and max-abs over pixels whose
    producing sites agree on the bucket (site_to_out maps a site-mask to an output mask)."""
    got = np.asarray(got, np.float64)
    err = np.abs(got - ref)
    fields = dict(max_abs=float(err.max()), psnr=psnr(got - ref), n_out=int(err.size))
    if got_rows is not None:
        gr = np.asarray(got_rows).reshape(-1)
        rr = np.asarray(ref_rows).reshape(-1)
        agree = gr == rr
        fields["bucket_agree"] = float(agree.mean())
        fields["bucket_disagree_n"] = int((~agree).sum())
        if site_to_out is not None:
            mask = site_to_out(agree.reshape(site_shape))
            fields["max_abs_agreeing"] = float(err[mask].max()) if mask.any() else 0.0
    report(case, **fields)
    return fields


def nnedi3_crop_ref(img, nns, wh, ox0, oy0, ow, oh, margin=8):
    """FP64 oracle output of the 2x NNEDI3 hook over the output region
    [ox0, ox0+ow) x [oy0, oy0+oh) of a 2W x 2H result, computed on an input crop with
    `margin` extra input pixels on every side that is not a frame edge.  A final pixel
    depends on input pixels within 5 rows and 7 columns of its position (double_y window
    rows -2..3 under double_x's rows -3..4; columns -3..4 under the x pass's -2..3), so
    8 suffices and the region equals the full-frame oracle's."""
    from oracle.nnedi3 import nnedi3
    H, W = img.shape
    x0, y0 = max(ox0 // 2 - margin, 0), max(oy0 // 2 - margin, 0)
    x1, y1 = min((ox0 + ow + 1) // 2 + margin, W), min((oy0 + oh + 1) // 2 + margin, H)
    out = nnedi3(np.asarray(img[y0:y1, x0:x1], np.float64), nns, wh)
    return out[oy0 - 2 * y0:oy0 - 2 * y0 + oh, ox0 - 2 * x0:ox0 - 2 * x0 + ow]
