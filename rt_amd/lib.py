"""ctypes binding of librt_hip.so (the C ABI of include/rt_hip.h).

Device memory and streams come from PyTorch-ROCm (plumbing only): buffers are
torch tensors on a HIP device whose data_ptr() is handed to the engine.  The
engine library is required: loading fails loudly when it is not built.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from . import scheme as S

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librt_hip.so")
_LIB = None

_dp = C.POINTER(C.c_double)
_vp = C.c_void_p


class EngineError(RuntimeError):
    pass


def lib():
    """Load librt_hip.so; raise if it is not built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise EngineError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.rt_hip_init.argtypes = [C.c_int]
        L.rt_hip_info.argtypes = [C.c_char_p, C.c_size_t]
        L.rt_hip_last_error.argtypes = [C.c_char_p, C.c_size_t]
        L.rt_hip_validate_scheme.argtypes = [_dp, C.c_int, _dp, C.c_size_t]
        L.rt_hip_scheme_create.argtypes = [_dp, C.c_int, _dp, C.c_size_t, C.POINTER(_vp)]
        L.rt_hip_scheme_destroy.argtypes = [_vp]
        L.rt_hip_trace.argtypes = ([_vp, _dp, C.c_uint64, C.c_size_t, C.c_size_t] + [_vp] * 8 +
                                   [C.c_int, _vp, _vp, _vp, _vp, _vp, _vp])
        L.rt_hip_set_jit.argtypes = [C.c_int]
        L.rt_hip_set_ltab.argtypes = [C.c_int]
        L.rt_hip_scheme_set_launch.argtypes = [_vp, C.c_int, C.c_int, C.c_int]
        L.rt_hip_scheme_get_launch.argtypes = [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.rt_hip_scheme_set_unroll.argtypes = [_vp, C.c_int]
        L.rt_hip_scheme_get_unroll.argtypes = [_vp, C.POINTER(C.c_int)]
        L.rt_hip_autotune.argtypes = [_vp, _dp, C.c_uint64, C.c_size_t, C.c_int, C.POINTER(C.c_double), _vp]
        L.rt_hip_last_kernel.argtypes = [C.c_char_p, C.c_size_t]
        L.rt_hip_batch_create.argtypes = [_dp, C.c_int, C.c_int, _dp, C.c_size_t, C.POINTER(C.c_uint64),
                                          C.POINTER(_vp)]
        L.rt_hip_batch_update.argtypes = [_vp, _dp, _dp, C.c_size_t]
        L.rt_hip_batch_destroy.argtypes = [_vp]
        L.rt_hip_batch_set_beams.argtypes = [_vp, _dp]
        L.rt_hip_batch_info.argtypes = [_vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.rt_hip_batch_set_launch.argtypes = [_vp, C.c_int, C.c_int, C.c_int]
        L.rt_hip_batch_trace.argtypes = [_vp, _dp, C.c_uint64, C.c_int] + [_vp] * 8 + [_vp, _vp, _vp]
        L.rt_hip_batch_graph_create.argtypes = ([_vp, _dp, C.c_uint64, C.c_int] + [_vp] * 8 +
                                                [_vp, _vp, _vp, C.POINTER(_vp)])
        L.rt_hip_adjoint.argtypes = [_vp, _dp, C.c_uint64, C.c_size_t, _vp, C.c_size_t,
                                     C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int, _vp, _vp, _vp]
        L.rt_hip_group_spots.argtypes = [_vp, _vp, _vp, C.c_size_t, C.c_size_t, _vp, _vp]
        L.rt_hip_group_spots_noalloc.argtypes = [_vp, _vp, _vp, C.c_size_t, C.c_size_t, _vp, _vp]
        L.rt_hip_trace_record.argtypes = [_vp, _dp, C.c_uint64, C.c_size_t] + [_vp] * 8 + [C.c_int, _vp, _vp,
                                                                                            _vp, _vp]
        L.rt_hip_rearrange.argtypes = [_vp, C.c_size_t, _vp, C.POINTER(C.c_uint64), _vp]
        L.rt_hip_beam_order.argtypes = [C.c_size_t, _vp, _vp, C.POINTER(C.c_uint8)] + [_vp] * 8 + [_vp]
        L.rt_hip_spot_loss.argtypes = [_vp, C.c_size_t, _vp, C.c_double, _vp, _vp, _vp]
        L.rt_hip_batch_group_spots.argtypes = [_vp, _vp, _vp, _vp, C.c_size_t, _vp, _vp]
        L.rt_hip_trace_spots.argtypes = [_vp, _dp, C.c_uint64, C.c_size_t, C.c_size_t, _vp, _vp, _vp]
        L.rt_hip_batch_trace_group_spots.argtypes = [_vp, _dp, C.c_uint64, C.c_size_t, _vp, _vp, _vp]
        L.rt_hip_batch_descend.argtypes = [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int, _vp, _vp, _vp]
        L.rt_hip_batch_adjoint.argtypes = [_vp, _dp, C.c_uint64, _vp, C.c_size_t, C.POINTER(C.c_int),
                                           C.POINTER(C.c_int), C.c_int, _vp, _vp, _vp]
        L.rt_hip_graph_launch.argtypes = [_vp, _vp]
        L.rt_hip_graph_destroy.argtypes = [_vp]
        L.rt_hip_jit_check.argtypes = [_dp, C.c_int, _dp, C.c_size_t, C.c_int, C.c_char_p, C.c_size_t,
                                       C.POINTER(C.c_size_t)]
        # per-stage entry points: host arrays (records, angles, beam descriptions) as
        # double*, device buffers and the stream as void*
        n, u64, lng, dbl = C.c_size_t, C.c_uint64, C.c_long, C.c_double
        L.rt_hip_stage_transform.argtypes = [_dp, _dp, n] + [_vp] * 6 + [_vp]
        L.rt_hip_stage_normalize.argtypes = [n] + [_vp] * 3 + [_vp]
        L.rt_hip_stage_intersect.argtypes = [_dp, n] + [_vp] * 6 + [_vp, _vp, _vp, _vp]
        L.rt_hip_stage_mask.argtypes = [_dp, n, _vp, _vp, _vp, _vp, _vp]
        L.rt_hip_stage_reflect.argtypes = [_dp, n] + [_vp] * 6 + [_vp]
        L.rt_hip_stage_refract.argtypes = [_dp, n] + [_vp] * 6 + [_vp, _vp, _vp, _vp]
        L.rt_hip_stage_diffract.argtypes = [_dp, n] + [_vp] * 6 + [_vp, _vp, _vp, _vp, _vp]
        L.rt_hip_beam_positions.argtypes = [_dp, u64, n, n, _vp, _vp, _vp, _vp]
        L.rt_hip_beam_directions.argtypes = [_dp, u64, n, n, _vp, _vp, _vp, _vp, _vp, _vp]
        L.rt_hip_beam_wavelengths.argtypes = [_dp, u64, n, n, _vp, _vp]
        L.rt_hip_ccd_coordinates.argtypes = [n, _vp, _vp, dbl, dbl, dbl, dbl, _vp, _vp, _vp]
        L.rt_hip_ccd_accumulate.argtypes = [n, _vp, _vp, _vp, n, n, _vp, _vp, _vp, _vp]
        L.rt_hip_ccd_flux.argtypes = [n, _vp, _vp, _vp, _dp, dbl, _dp, dbl, dbl, dbl, dbl, dbl, n, n,
                                      _vp, _vp, _vp]
        L.rt_hip_scale.argtypes = [n, _vp, _vp, _vp]
        L.rt_hip_grating_energy.argtypes = [n, _vp, lng, dbl, dbl, dbl, dbl, dbl, dbl, _vp, _vp]
        L.rt_hip_spline_init.argtypes = [n, _dp, _dp, _dp]
        L.rt_hip_spline_eval.argtypes = [_dp, n, _vp, _vp, _vp, _vp]
        _LIB = L
    return _LIB


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().rt_hip_last_error(buf, len(buf))
    return buf.value.decode()


def check(rc: int, what: str):
    if rc != 0:
        raise EngineError(f"{what} failed with status {rc}: {last_error()}")


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _tp(t) -> Optional[int]:
    """device address of a torch tensor (or None)."""
    return None if t is None else t.data_ptr()


def info() -> str:
    buf = C.create_string_buffer(4096)
    check(lib().rt_hip_info(buf, len(buf)), "rt_hip_info")
    return buf.value.decode()


def validate(scheme: "S.Scheme") -> int:
    t = np.ascontiguousarray(scheme.table)
    p = np.ascontiguousarray(scheme.pool)
    return lib().rt_hip_validate_scheme(_dptr(t), scheme.n_surf, _dptr(p), len(p))


class DeviceScheme:
    """A scheme uploaded to the current HIP device (rt_hip_scheme)."""

    def __init__(self, scheme: "S.Scheme"):
        self.scheme = scheme
        self._t = np.ascontiguousarray(scheme.table, dtype=np.float64)
        self._p = np.ascontiguousarray(scheme.pool, dtype=np.float64)
        h = _vp()
        check(lib().rt_hip_scheme_create(_dptr(self._t), scheme.n_surf, _dptr(self._p), len(self._p),
                                          C.byref(h)), "rt_hip_scheme_create")
        self.handle = h

    def close(self):
        if self.handle:
            lib().rt_hip_scheme_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def trace(self, rays=None, flag=None, *, seed=0, begin=0, n=None, flags=0, beam=None,
              hist=None, flux=None, stats=None, dump=None, dump_flag=None, stream=None):
        """Launch rt_hip_trace.  `rays` is a (7, n) float64 CUDA tensor or None;
        with rays given and RT_TRACE_GENERATE in flags they are outputs."""
        import torch
        b = np.ascontiguousarray((beam if beam is not None else self.scheme.beam).pack())
        if n is None:
            n = rays.shape[1] if rays is not None else self.scheme.beam.n_rays
        ptrs = [None] * 7
        if rays is not None:
            assert rays.dtype == torch.float64 and rays.is_cuda and rays.is_contiguous()
            assert rays.shape[0] == 7 and rays.shape[1] >= n
            ptrs = [rays[k].data_ptr() for k in range(7)]
        if flag is not None:
            assert flag.dtype == torch.int8 and flag.is_cuda and flag.numel() >= n
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        rc = lib().rt_hip_trace(self.handle, _dptr(b), seed, begin, n, *ptrs, _tp(flag), flags,
                                _tp(hist), _tp(flux), _tp(stats), _tp(dump), _tp(dump_flag), stream)
        check(rc, "rt_hip_trace")

    def trace_record(self, rays, flag, loss, *, seed=0, n=None, extents=None, stats=None, beam=None,
                     stream=None):
        """rt_hip_trace_record: the generated beam of `seed` traced into rays
        (7 x n), flag, loss (n uint8 CUDA tensor) and, when given, extents
        (n_surf x RT_EXT_PER_SURF float64 CUDA tensor)."""
        import torch
        b = np.ascontiguousarray((beam if beam is not None else self.scheme.beam).pack())
        if n is None:
            n = rays.shape[1]
        assert rays.dtype == torch.float64 and rays.is_cuda and rays.is_contiguous() and rays.shape[1] >= n
        assert flag.dtype == torch.int8 and loss.dtype == torch.uint8 and loss.numel() >= n
        if extents is not None:
            assert extents.dtype == torch.float64 and extents.is_contiguous()
            assert extents.numel() >= self.scheme.n_surf * EXT_PER_SURF
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        rc = lib().rt_hip_trace_record(self.handle, _dptr(b), seed, n, *[rays[k].data_ptr() for k in range(7)],
                                       _tp(flag), S.RT_TRACE_GENERATE | S.RT_TRACE_STORE, _tp(stats), _tp(loss),
                                       _tp(extents), stream)
        check(rc, "rt_hip_trace_record")

    def trace_spots(self, n_groups: int, *, seed=0, n=None, stats=None, beam=None, stream=None, out=None):
        """rt_hip_trace_spots: the generated beam of `seed` traced with the
        (n_groups, SPOT_STRIDE) float64 CUDA spots of the rays alive at the end
        formed in the trace (no rays stored)."""
        import torch
        b = np.ascontiguousarray((beam if beam is not None else self.scheme.beam).pack())
        if n is None:
            n = self.scheme.beam.n_rays
        if out is None:
            out = torch.empty((n_groups, SPOT_STRIDE), dtype=torch.float64, device="cuda")
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        check(lib().rt_hip_trace_spots(self.handle, _dptr(b), seed, n, n_groups, _tp(out), _tp(stats), stream),
              "rt_hip_trace_spots")
        return out

    def adjoint(self, group_seeds, param_surface, param_field, *, seed=0, n=None, grad=None, stats=None,
                beam=None, stream=None):
        """rt_hip_adjoint: dLoss/dparameters for a loss seeded per ray group by
        group_seeds (CUDA float64 tensor, n_groups x 3: ax, ay, b).  Returns grad
        (CUDA float64 tensor, one value per parameter)."""
        import torch
        b = np.ascontiguousarray((beam if beam is not None else self.scheme.beam).pack())
        if n is None:
            n = self.scheme.beam.n_rays
        ps = np.ascontiguousarray(param_surface, dtype=np.int32)
        pf = np.ascontiguousarray(param_field, dtype=np.int32)
        assert len(ps) == len(pf)
        assert group_seeds.dtype == torch.float64 and group_seeds.is_cuda and group_seeds.is_contiguous()
        if grad is None:
            grad = torch.zeros(len(ps), dtype=torch.float64, device=group_seeds.device)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        ip = C.POINTER(C.c_int)
        check(lib().rt_hip_adjoint(self.handle, _dptr(b), seed, n, group_seeds.data_ptr(), group_seeds.shape[0],
                                   ps.ctypes.data_as(ip), pf.ctypes.data_as(ip), len(ps), _tp(grad), _tp(stats),
                                   stream), "rt_hip_adjoint")
        return grad

    def set_launch(self, block: int = 256, waves: int = 7, grid_mult: int = 16):
        """Launch of the generated kernel (rt_hip_scheme_set_launch)."""
        check(lib().rt_hip_scheme_set_launch(self.handle, block, waves, grid_mult), "rt_hip_scheme_set_launch")

    def set_unroll(self, full: bool):
        """Surface unroll depth (rt_hip_scheme_set_unroll): True, the generated
        kernel with the surface chain unrolled; False, the generic runtime loop."""
        check(lib().rt_hip_scheme_set_unroll(self.handle, 1 if full else 0), "rt_hip_scheme_set_unroll")

    def get_unroll(self) -> bool:
        u = C.c_int()
        check(lib().rt_hip_scheme_get_unroll(self.handle, C.byref(u)), "rt_hip_scheme_get_unroll")
        return bool(u.value)

    def get_launch(self):
        b, w, m = C.c_int(), C.c_int(), C.c_int()
        check(lib().rt_hip_scheme_get_launch(self.handle, C.byref(b), C.byref(w), C.byref(m)),
              "rt_hip_scheme_get_launch")
        return b.value, w.value, m.value

    def autotune(self, n_sample: int, flags: int, seed: int = 0, beam=None, stream=None) -> float:
        """Time the candidate launches on n_sample generated rays and keep the
        fastest (rt_hip_autotune); returns its time in ms."""
        import torch
        b = np.ascontiguousarray((beam if beam is not None else self.scheme.beam).pack())
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        ms = C.c_double(0.0)
        check(lib().rt_hip_autotune(self.handle, _dptr(b), seed, n_sample, flags, C.byref(ms), stream),
              "rt_hip_autotune")
        return ms.value


def set_jit(on: bool) -> bool:
    """Use kernels generated per scheme structure (True) or the generic kernel."""
    return bool(lib().rt_hip_set_jit(1 if on else 0))


def set_ltab(on: bool) -> bool:
    """Generated traces of grid-wavelength beams read the per-wavelength table
    of spline values (True, the default) or evaluate the splines per ray."""
    return bool(lib().rt_hip_set_ltab(1 if on else 0))


def last_kernel() -> str:
    buf = C.create_string_buffer(64)
    lib().rt_hip_last_kernel(buf, len(buf))
    return buf.value.decode()


def jit_check(scheme: "S.Scheme", beam_mode: int):
    """Generate and compile (without a device) the kernel of `scheme`;
    returns (status, code bytes, source + compiler log)."""
    t = np.ascontiguousarray(scheme.table)
    p = np.ascontiguousarray(scheme.pool)
    buf = C.create_string_buffer(1 << 17)
    size = C.c_size_t(0)
    rc = lib().rt_hip_jit_check(_dptr(t), scheme.n_surf, _dptr(p), len(p), beam_mode, buf, len(buf),
                                C.byref(size))
    return rc, size.value, buf.value.decode()


def init(device: int = 0):
    check(lib().rt_hip_init(device), "rt_hip_init")


SPOT_STRIDE = 6     # RT_SPOT_STRIDE: n, centroid x, y, variance x, y, RMS radius
EXT_PER_SURF = 12   # RT_EXT_PER_SURF


def rearrange(loss, stream=None):
    """rt_hip_rearrange: (perm, counts): the ray at each position of the
    saved beam (uint64 CUDA tensor, int64 dtype) and the rays of every loss
    code (numpy uint64[256])."""
    import torch
    n = loss.numel()
    perm = torch.empty(n, dtype=torch.int64, device=loss.device)
    counts = np.zeros(256, dtype=np.uint64)
    st = torch.cuda.current_stream().cuda_stream if stream is None else stream
    check(lib().rt_hip_rearrange(_tp(loss), n, _tp(perm), counts.ctypes.data_as(C.POINTER(C.c_uint64)),
                                 C.c_void_p(st)), "rt_hip_rearrange")
    return perm, counts


def beam_order(perm, loss, renorm, rays, stream=None):
    """rt_hip_beam_order: the 7 x n arrays of Beam::Save in file order (CUDA
    float64 tensor); renorm: numpy uint8[256]."""
    import torch
    n = perm.numel()
    out = torch.empty((7, n), dtype=torch.float64, device=rays.device)
    ren = np.ascontiguousarray(renorm, dtype=np.uint8)
    st = torch.cuda.current_stream().cuda_stream if stream is None else stream
    check(lib().rt_hip_beam_order(n, _tp(perm), _tp(loss), ren.ctypes.data_as(C.POINTER(C.c_uint8)),
                                  *[rays[k].data_ptr() for k in range(7)], _tp(out), C.c_void_p(st)),
          "rt_hip_beam_order")
    return out


def group_spots(rays, flag, n_groups: int, stream=None, out=None, noalloc=False):
    """rt_hip_group_spots: (n_groups, SPOT_STRIDE) float64 CUDA tensor of the
    spots of the rays with flag 1, ray r in group r mod n_groups (rays: the
    7 x n state on the device, flag: n int8).  noalloc: through
    rt_hip_group_spots_noalloc (static scratch, bounded size)."""
    import torch
    n = rays.shape[1]
    if rays.dtype != torch.float64 or not rays.is_contiguous() or flag.dtype != torch.int8:
        raise EngineError("rays must be contiguous float64 and flag int8")
    if out is None:
        out = torch.empty((n_groups, SPOT_STRIDE), dtype=torch.float64, device=rays.device)
    st = torch.cuda.current_stream().cuda_stream if stream is None else stream
    name = "rt_hip_group_spots_noalloc" if noalloc else "rt_hip_group_spots"
    check(getattr(lib(), name)(C.c_void_p(rays[0].data_ptr()), C.c_void_p(rays[1].data_ptr()),
                               C.c_void_p(flag.data_ptr()), n, n_groups, C.c_void_p(out.data_ptr()),
                               C.c_void_p(st)), name)
    return out


def spot_loss(spots, target=None, weight: float = 0.0, stream=None, seeds=None, loss=None):
    """rt_hip_spot_loss: (loss, seeds) as CUDA tensors (a 1-element tensor and
    n_groups x 3) from group_spots' output; target: n_groups x 2 or None."""
    import torch
    G = spots.shape[0]
    if seeds is None:
        seeds = torch.empty((G, 3), dtype=torch.float64, device=spots.device)
    if loss is None:
        loss = torch.empty(1, dtype=torch.float64, device=spots.device)
    st = torch.cuda.current_stream().cuda_stream if stream is None else stream
    tp = None if target is None else C.c_void_p(target.data_ptr())
    check(lib().rt_hip_spot_loss(C.c_void_p(spots.data_ptr()), G, tp, float(weight), C.c_void_p(seeds.data_ptr()),
                                 C.c_void_p(loss.data_ptr()), C.c_void_p(st)), "rt_hip_spot_loss")
    return loss, seeds


def pack_batch(instances):
    """(n_inst x n_surf x RT_SURF_STRIDE tables, shared pool) of schemes of one structure"""
    pool = np.ascontiguousarray(instances[0].pool, dtype=np.float64)
    for s in instances[1:]:
        if not np.array_equal(s.pool, pool):
            raise ValueError("batched instances must share one spline pool")
    tables = np.ascontiguousarray(np.stack([np.asarray(s.table, dtype=np.float64) for s in instances]))
    return tables, pool


class DeviceBatch:
    """Instances of one scheme structure on the device (rt_hip_batch)."""

    def __init__(self, instances, ray_counts):
        self.instances = instances
        self.n_inst = len(instances)
        self.n_surf = instances[0].n_surf
        self._t, self._p = pack_batch(instances)
        self.counts = np.ascontiguousarray(ray_counts, dtype=np.uint64)
        h = _vp()
        check(lib().rt_hip_batch_create(_dptr(self._t), self.n_inst, self.n_surf, _dptr(self._p), len(self._p),
                                        self.counts.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(h)),
              "rt_hip_batch_create")
        self.handle = h
        self.total = int(self.counts.sum())

    def close(self):
        if self.handle:
            lib().rt_hip_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def update(self, instances):
        t, p = pack_batch(instances)
        self._t, self._p = t, p
        check(lib().rt_hip_batch_update(self.handle, _dptr(t), _dptr(p), len(p)), "rt_hip_batch_update")

    def set_launch(self, block=256, waves=7, grid_mult=16):
        check(lib().rt_hip_batch_set_launch(self.handle, block, waves, grid_mult), "rt_hip_batch_set_launch")

    def info(self):
        n, t = C.c_uint64(), C.c_uint64()
        check(lib().rt_hip_batch_info(self.handle, C.byref(n), C.byref(t)), "rt_hip_batch_info")
        return n.value, t.value

    def set_beams(self, beams):
        """rt_hip_batch_set_beams: a Beam per instance (one beam type), or None
        for the first instance's beam everywhere."""
        if beams is None:
            self._beams = None
            check(lib().rt_hip_batch_set_beams(self.handle, None), "rt_hip_batch_set_beams")
            return
        assert len(beams) == self.n_inst
        self._beams = np.ascontiguousarray(np.concatenate([b.pack() for b in beams]))
        check(lib().rt_hip_batch_set_beams(self.handle, _dptr(self._beams)), "rt_hip_batch_set_beams")

    def _args(self, beam, rays, flag):
        b = np.ascontiguousarray((beam if beam is not None else self.instances[0].beam).pack())
        ptrs = [None] * 8
        if rays is not None:
            assert rays.shape == (7, self.total) and rays.is_contiguous()
            ptrs = [rays[k].data_ptr() for k in range(7)] + [_tp(flag)]
        return b, ptrs

    def trace(self, *, seed=0, flags=0, rays=None, flag=None, spots=None, stats=None, beam=None, stream=None):
        """rt_hip_batch_trace; rays (7, total) and flag are filled with RT_TRACE_STORE,
        spots (n_inst, SPOT_STRIDE) float64 receives the per-instance spots."""
        import torch
        b, ptrs = self._args(beam, rays, flag)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        self._beam = b
        check(lib().rt_hip_batch_trace(self.handle, _dptr(b), seed, flags, *ptrs, _tp(spots), _tp(stats), stream),
              "rt_hip_batch_trace")

    def trace_group_spots(self, n_groups: int, *, seed=0, stats=None, beam=None, stream=None, out=None):
        """rt_hip_batch_trace_group_spots: the batched trace of generated beams
        with the (n_inst, n_groups, SPOT_STRIDE) float64 CUDA spots formed in
        the trace (no rays stored)."""
        import torch
        b, _ = self._args(beam, None, None)
        if out is None:
            out = torch.empty((self.n_inst, n_groups, SPOT_STRIDE), dtype=torch.float64, device="cuda")
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        self._beam = b
        check(lib().rt_hip_batch_trace_group_spots(self.handle, _dptr(b), seed, n_groups, _tp(out), _tp(stats),
                                                   stream), "rt_hip_batch_trace_group_spots")
        return out

    def group_spots(self, rays, flag, n_groups: int, stream=None, out=None):
        """rt_hip_batch_group_spots: (n_inst, n_groups, SPOT_STRIDE) float64 CUDA
        tensor from the stored rays (7, total) of a batched trace."""
        import torch
        assert rays.shape == (7, self.total) and rays.is_contiguous() and flag.dtype == torch.int8
        if out is None:
            out = torch.empty((self.n_inst, n_groups, SPOT_STRIDE), dtype=torch.float64, device=rays.device)
        st = torch.cuda.current_stream().cuda_stream if stream is None else stream
        check(lib().rt_hip_batch_group_spots(self.handle, C.c_void_p(rays[0].data_ptr()),
                                             C.c_void_p(rays[1].data_ptr()), C.c_void_p(flag.data_ptr()), n_groups,
                                             C.c_void_p(out.data_ptr()), C.c_void_p(st)), "rt_hip_batch_group_spots")
        return out

    def adjoint(self, group_seeds, param_surface, param_field, *, seed=0, grad=None, stats=None, beam=None,
                stream=None):
        """rt_hip_batch_adjoint: per-instance gradients (n_inst, n_params) for
        the loss seeded per (instance, group) by group_seeds (CUDA float64,
        n_inst x n_groups x 3)."""
        import torch
        b = np.ascontiguousarray((beam if beam is not None else self.instances[0].beam).pack())
        ps = np.ascontiguousarray(param_surface, dtype=np.int32)
        pf = np.ascontiguousarray(param_field, dtype=np.int32)
        assert len(ps) == len(pf)
        assert group_seeds.dtype == torch.float64 and group_seeds.is_cuda and group_seeds.is_contiguous()
        assert group_seeds.shape[0] == self.n_inst and group_seeds.shape[2] == 3
        if grad is None:
            grad = torch.zeros((self.n_inst, len(ps)), dtype=torch.float64, device=group_seeds.device)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        ip = C.POINTER(C.c_int)
        self._adj_keep = (b, ps, pf)
        check(lib().rt_hip_batch_adjoint(self.handle, _dptr(b), seed, group_seeds.data_ptr(), group_seeds.shape[1],
                                         ps.ctypes.data_as(ip), pf.ctypes.data_as(ip), len(ps), _tp(grad),
                                         _tp(stats), stream), "rt_hip_batch_adjoint")
        return grad

    def descend(self, param_surface, param_field, grad, lr, stream=None):
        """rt_hip_batch_descend: field -= lr[j] * grad[i][j] on the device tables
        (grad: CUDA float64 n_inst x n_params, lr: CUDA float64 n_params)."""
        import torch
        ps = np.ascontiguousarray(param_surface, dtype=np.int32)
        pf = np.ascontiguousarray(param_field, dtype=np.int32)
        st = torch.cuda.current_stream().cuda_stream if stream is None else stream
        ip = C.POINTER(C.c_int)
        self._desc_keep = (ps, pf)
        check(lib().rt_hip_batch_descend(self.handle, ps.ctypes.data_as(ip), pf.ctypes.data_as(ip), len(ps),
                                         C.c_void_p(grad.data_ptr()), C.c_void_p(lr.data_ptr()), C.c_void_p(st)),
              "rt_hip_batch_descend")

    def graph(self, *, stream, seed=0, flags=0, rays=None, flag=None, spots=None, stats=None, beam=None):
        """Record the batched trace as a hipGraph on `stream` (a torch.cuda.Stream)."""
        b, ptrs = self._args(beam, rays, flag)
        g = _vp()
        check(lib().rt_hip_batch_graph_create(self.handle, _dptr(b), seed, flags, *ptrs, _tp(spots), _tp(stats),
                                              stream.cuda_stream, C.byref(g)), "rt_hip_batch_graph_create")
        return Graph(g, b)


class Graph:
    def __init__(self, handle, keep):
        self.handle, self._keep = handle, keep

    def launch(self, stream):
        check(lib().rt_hip_graph_launch(self.handle, stream.cuda_stream), "rt_hip_graph_launch")

    def close(self):
        if self.handle:
            lib().rt_hip_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- librt_scheme_hip.so: the Scheme path (XML loader, Scheme::run,
# ComputeQE, ccd_frame)

SCHEME_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "librt_scheme_hip.so")
_SCHEME_LIB = None


def scheme_lib():
    """Load librt_scheme_hip.so; raise if it is not built."""
    global _SCHEME_LIB
    if _SCHEME_LIB is None:
        if not os.path.exists(SCHEME_LIB_PATH):
            raise EngineError(f"{SCHEME_LIB_PATH} is not built; run __graft_entry__.build()")
        lib()                                     # librt_hip.so first (the scheme library links it)
        L = C.CDLL(SCHEME_LIB_PATH)
        cp = C.c_char_p
        L.rt_scheme_run_file.argtypes = [cp, cp, cp]
        L.rt_scheme_compute_qe.argtypes = [cp, C.c_size_t, cp]
        L.rt_scheme_pack_file.argtypes = [cp, _dp, C.c_int, C.POINTER(C.c_int), _dp, C.c_size_t,
                                          C.POINTER(C.c_size_t), _dp, C.POINTER(C.c_uint64)]
        L.rt_ccd_frame.argtypes = [cp, cp, cp, cp, cp]
        L.rt_write_fits_image.argtypes = [cp, C.c_size_t, C.c_size_t, _dp, C.c_int, C.POINTER(cp),
                                          C.POINTER(cp)]
        _SCHEME_LIB = L
    return _SCHEME_LIB


def _b(s):
    return None if s is None else os.fsencode(s)


def scheme_pack_file(path: str, max_surf: int = 32, max_pool: int = 1 << 20):
    """rt_scheme_pack_file: (status, table, pool, beam_desc, n_rays) of an XML
    scheme as Scheme::run would hand it to the engine (no device used)."""
    table = np.zeros(max_surf * S.SURF_STRIDE)
    pool = np.zeros(max_pool)
    beam = np.zeros(S.BEAM_STRIDE)
    ns, pl, nr = C.c_int(0), C.c_size_t(0), C.c_uint64(0)
    rc = scheme_lib().rt_scheme_pack_file(_b(path), _dptr(table), max_surf, C.byref(ns), _dptr(pool), max_pool,
                                          C.byref(pl), _dptr(beam), C.byref(nr))
    if rc:
        return rc, None, None, None, 0
    return 0, table[:ns.value * S.SURF_STRIDE].copy(), pool[:pl.value].copy(), beam, nr.value


def scheme_run_file(path: str, list_file: str, log_file: str = None) -> int:
    """rt_scheme_run_file: ray-tracing/main.cpp's flow on the engine."""
    return scheme_lib().rt_scheme_run_file(_b(path), _b(list_file), _b(log_file))


def scheme_compute_qe(path: str, nlambda: int, qe_file: str) -> int:
    """rt_scheme_compute_qe: the qe_comp flow (Scheme::ComputeQE)."""
    return scheme_lib().rt_scheme_compute_qe(_b(path), nlambda, _b(qe_file))


def ccd_frame(rt_list: str, qe_file: str, spec_file: str, ccd_spec: str, fits_file: str) -> int:
    """rt_ccd_frame: ccd_frame's program on the engine (exit code)."""
    return scheme_lib().rt_ccd_frame(_b(rt_list), _b(qe_file), _b(spec_file), _b(ccd_spec), _b(fits_file))


def write_fits_image(path: str, image: np.ndarray, keys=()):
    """rt_write_fits_image of a (ydim, xdim) float64 image"""
    img = np.ascontiguousarray(image, dtype=np.float64)
    ks = (C.c_char_p * max(1, len(keys)))(*[k.encode() for k, _ in keys])
    vs = (C.c_char_p * max(1, len(keys)))(*[v.encode() for _, v in keys])
    return scheme_lib().rt_write_fits_image(_b(path), img.shape[1], img.shape[0], _dptr(img), len(keys), ks, vs)
