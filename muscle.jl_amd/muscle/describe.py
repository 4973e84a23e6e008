"""Host-only inspection of how a contraction is lowered (muscle_hip_describe)."""
from __future__ import annotations

import ctypes

from . import _lib

CLASS_NAMES = {0: "batch", 1: "left", 2: "right", 3: "contract", 4: "sum_a", 5: "sum_b"}


def describe(inds_c, shape_c, inds_a, shape_a, inds_b, shape_b, dtypes=("c64", "c64", "c64"),
             strides=(None, None, None), precision: str = "native", target_rel_error: float = 1e-12) -> dict:
    """Classify every index and report the fused GEMM shape and kernel choice.

    Index labels may be any hashables; shapes are per-index extents; strides
    default to column-major.  dtypes are "f32", "f64", "c32", "c64".
    """
    code = {"f32": _lib.F32, "f64": _lib.F64, "c32": _lib.C32, "c64": _lib.C64}
    labels = {}
    for i in list(inds_c) + list(inds_a) + list(inds_b):
        labels.setdefault(i, len(labels) + 1)
    descs = []
    for inds, shape, dt, st in zip((inds_c, inds_a, inds_b), (shape_c, shape_a, shape_b), dtypes, strides):
        if st is None:
            s, st = 1, []
            for n in shape:
                st.append(s)
                s *= max(int(n), 1)
        descs.append(_lib.Desc(code[dt], 0, [labels[i] for i in inds], list(shape), list(st)))
    info = _lib.Info()
    nmax = max(len(labels), 1)
    modes = (ctypes.c_int32 * nmax)()
    classes = (ctypes.c_int32 * nmax)()
    nm = ctypes.c_int32()
    _lib.check(_lib.lib.muscle_hip_describe(descs[0].ref, descs[1].ref, descs[2].ref,
                                            _lib.PRECISION[precision], float(target_rel_error),
                                            ctypes.byref(info), modes, classes,
                                            ctypes.byref(nm)))
    inv = {v: k for k, v in labels.items()}
    out = info.as_dict()
    out["classes"] = {inv[modes[i]]: CLASS_NAMES[classes[i]] for i in range(nm.value)}
    out["kernel_name"] = {0: "none", 1: "gemm_mfma", 2: "reduce", 3: "split", 4: "elementwise", 5: "unused",
                          6: "permute", 7: "split_crt", 8: "split_crt_gaussian"}[info.kernel]
    return out
