"""Python mirror of the reference's Matrix_Format plugin contract over the
gfx950 C ABI (include/spmvhip.h).

    fmt = csr_to_format(row_ptr, col_ind, values, m, n, nnz, symmetric=0,
                        symmetry_expanded=0, kind="csr_vec")
    fmt.spmv(x, y)                       # host arrays, y written in place
    fmt.spmv_device(dx, dy)              # DeviceArray operands, asynchronous

as src/spmv_kernels/spmv_kernel.h:8-29 declares them (csr_to_format,
Matrix_Format::spmv / statistics_start / statistics_print_data,
statistics_print_labels). Errors raise SpmvHipError carrying the C status,
where the reference would exit().
"""
import ctypes

import numpy as np

from . import _native

_lib = _native.load("spmvhip")

KINDS = {"csr_naive": 0, "csr_vec": 1, "sell": 2, "sell_c_s": 3, "div_rf": 4, "sell_sorted": 5}
KIND_NAMES = {v: k for k, v in KINDS.items()}


class Params(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("C", ctypes.c_int), ("sigma", ctypes.c_long),
                ("lanes_per_row", ctypes.c_int), ("block_nnz", ctypes.c_int),
                ("split_nnz", ctypes.c_int), ("xcd_remap", ctypes.c_int),
                ("x_window", ctypes.c_int), ("unroll", ctypes.c_int),
                ("div_threads", ctypes.c_int), ("div_lanes", ctypes.c_int),
                ("div_packet_vals", ctypes.c_long), ("pipeline", ctypes.c_int),
                ("y_mode", ctypes.c_int), ("div_acc_copies", ctypes.c_int), ("narrow_loads", ctypes.c_int),
                ("div_layout", ctypes.c_int), ("div_reorder", ctypes.c_int),
                ("div_tolerance", ctypes.c_double), ("div_encoder", ctypes.c_int),
                ("reserved", ctypes.c_int * 1)]


class Info(ctypes.Structure):
    _fields_ = [("format_name", ctypes.c_char_p), ("m", ctypes.c_long), ("n", ctypes.c_long),
                ("nnz", ctypes.c_long), ("mem_footprint", ctypes.c_double),
                ("csr_mem_footprint", ctypes.c_double), ("conv_ms", ctypes.c_double),
                ("padded_nnz", ctypes.c_long), ("num_work_items", ctypes.c_long),
                ("num_split_rows", ctypes.c_long)]


_vp = ctypes.c_void_p
_sig = {
    "spmvhip_init": ([ctypes.c_int], ctypes.c_int),
    "spmvhip_device_name": ([ctypes.c_char_p, ctypes.c_long], ctypes.c_int),
    "spmvhip_malloc": ([ctypes.POINTER(_vp), ctypes.c_size_t], ctypes.c_int),
    "spmvhip_free": ([_vp], ctypes.c_int),
    "spmvhip_memcpy_h2d": ([_vp, _vp, ctypes.c_size_t], ctypes.c_int),
    "spmvhip_memcpy_d2h": ([_vp, _vp, ctypes.c_size_t], ctypes.c_int),
    "spmvhip_memset": ([_vp, ctypes.c_int, ctypes.c_size_t], ctypes.c_int),
    "spmvhip_synchronize": ([], ctypes.c_int),
    "spmvhip_last_error": ([], ctypes.c_char_p),
    "spmvhip_csr_to_format": ([_vp, _vp, _vp, ctypes.c_long, ctypes.c_long, ctypes.c_long,
                               ctypes.c_long, ctypes.c_long, ctypes.POINTER(Params),
                               ctypes.POINTER(_vp)], ctypes.c_int),
    "spmvhip_csr_to_format_device": ([_vp, _vp, _vp, ctypes.c_long, ctypes.c_long, ctypes.c_long,
                                      ctypes.POINTER(Params), ctypes.POINTER(_vp)], ctypes.c_int),
    "spmvhip_spmv": ([_vp, _vp, _vp], ctypes.c_int),
    "spmvhip_spmv_device": ([_vp, _vp, _vp, _vp], ctypes.c_int),
    "spmvhip_bind_host_xy": ([_vp, _vp, _vp], ctypes.c_int),
    "spmvhip_sync_host_y": ([_vp], ctypes.c_int),
    "spmvhip_statistics_start": ([_vp], ctypes.c_int),
    "spmvhip_statistics_print_data": ([_vp, ctypes.c_char_p, ctypes.c_long], ctypes.c_int),
    "spmvhip_statistics_print_labels": ([ctypes.c_char_p, ctypes.c_long], ctypes.c_int),
    "spmvhip_get_info": ([_vp, ctypes.POINTER(Info)], ctypes.c_int),
    "spmvhip_time_spmv": ([_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp,
                           ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
    "spmvhip_export": ([_vp, ctypes.c_int, _vp, ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
    "spmvhip_destroy": ([_vp], None),
    "spmvhip_batch_create": ([ctypes.c_int, ctypes.POINTER(_vp)], ctypes.c_int),
    "spmvhip_batch_set_queues": ([_vp, ctypes.c_int], ctypes.c_int),
    "spmvhip_batch_set_grid": ([_vp, ctypes.c_int], ctypes.c_int),
    "spmvhip_batch_set_xcd_ranges": ([_vp, ctypes.c_int], ctypes.c_int),
    "spmvhip_batch_set_solo_nnz": ([_vp, ctypes.c_long], ctypes.c_int),
    "spmvhip_batch_add": ([_vp, _vp, _vp, _vp], ctypes.c_int),
    "spmvhip_batch_finalize": ([_vp, ctypes.c_int], ctypes.c_int),
    "spmvhip_batch_run": ([_vp], ctypes.c_int),
    "spmvhip_batch_synchronize": ([_vp], ctypes.c_int),
    "spmvhip_batch_time": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp], ctypes.c_int),
    "spmvhip_batch_info": ([_vp, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long),
                            ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_long)], ctypes.c_int),
    "spmvhip_batch_destroy": ([_vp], None),
    "spmvhip_csr_stats_device": ([_vp, _vp, ctypes.c_long, ctypes.c_long, _vp], ctypes.c_int),
    "spmvhip_csr_stats_device_noalloc": ([_vp, _vp, ctypes.c_long, ctypes.c_long, _vp, _vp], ctypes.c_int),
    "spmvhip_bandwidth_probe": ([ctypes.c_int, _vp, _vp, ctypes.c_long, ctypes.c_int, _vp], ctypes.c_int),
    "spmvhip_gather_probe": ([ctypes.c_int, _vp, ctypes.c_long, ctypes.c_int, ctypes.c_double, _vp, _vp],
                             ctypes.c_int),
    "spmvhip_mlp_create": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint,
                            ctypes.POINTER(_vp)], ctypes.c_int),
    "spmvhip_mlp_num_params": ([_vp], ctypes.c_long),
    "spmvhip_mlp_set_params": ([_vp, _vp, ctypes.c_long], ctypes.c_int),
    "spmvhip_mlp_get_params": ([_vp, _vp, ctypes.c_long], ctypes.c_int),
    "spmvhip_mlp_predict": ([_vp, _vp, ctypes.c_int, _vp], ctypes.c_int),
    "spmvhip_mlp_gradients": ([_vp, _vp, _vp, _vp, ctypes.c_int, _vp, _vp], ctypes.c_int),
    "spmvhip_mlp_train": ([_vp, _vp, _vp, _vp, ctypes.c_long, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                           _vp], ctypes.c_int),
    "spmvhip_mlp_adam_step": ([_vp, _vp, ctypes.c_float], ctypes.c_int),
    "spmvhip_mlp_get_adam_state": ([_vp, _vp, _vp, ctypes.c_long, ctypes.POINTER(ctypes.c_long)], ctypes.c_int),
    "spmvhip_mlp_set_adam_state": ([_vp, _vp, _vp, ctypes.c_long, ctypes.c_long], ctypes.c_int),
    "spmvhip_mlp_destroy": ([_vp], None),
    "spmvhip_div_encode": ([_vp, _vp, _vp, ctypes.c_long, ctypes.c_long, ctypes.c_long, ctypes.c_int,
                            ctypes.c_int, ctypes.c_long, _vp, ctypes.POINTER(ctypes.c_size_t), _vp,
                            ctypes.POINTER(ctypes.c_long), _vp, ctypes.POINTER(ctypes.c_long)], ctypes.c_int),
    "spmvhip_div_reduce_precision": ([_vp, ctypes.c_long, ctypes.c_double, _vp], ctypes.c_int),
    "spmvhip_div_group_order": ([_vp, _vp, _vp, ctypes.c_long, ctypes.c_long, ctypes.c_long, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_long, _vp, ctypes.POINTER(ctypes.c_long)], ctypes.c_int),
    "spmvhip_div_device_encode": ([_vp, _vp, _vp, ctypes.c_long, ctypes.c_long, ctypes.c_long, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_long, ctypes.c_int, _vp, ctypes.POINTER(ctypes.c_size_t),
                                   _vp, ctypes.POINTER(ctypes.c_long)], ctypes.c_int),
}
for _name, (_args, _res) in _sig.items():
    _f = getattr(_lib, _name)
    _f.argtypes = _args
    _f.restype = _res

EXPORTED = tuple(_sig)


class SpmvHipError(RuntimeError):
    def __init__(self, code, where):
        msg = _lib.spmvhip_last_error()
        super().__init__("%s failed (%d): %s" % (where, code, msg.decode() if msg else ""))
        self.code = code


def _check(rc, where):
    if rc != 0:
        raise SpmvHipError(rc, where)


def init(device=0):
    _check(_lib.spmvhip_init(device), "spmvhip_init")


def device_name():
    buf = ctypes.create_string_buffer(256)
    _check(_lib.spmvhip_device_name(buf, 256), "spmvhip_device_name")
    return buf.value.decode()


def synchronize():
    _check(_lib.spmvhip_synchronize(), "spmvhip_synchronize")


class DeviceArray:
    """A flat device allocation holding a numpy dtype (plumbing only)."""

    def __init__(self, size, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.size = int(size)
        self.nbytes = self.size * self.dtype.itemsize
        p = _vp()
        _check(_lib.spmvhip_malloc(ctypes.byref(p), max(self.nbytes, 8)), "spmvhip_malloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.size
        _check(_lib.spmvhip_memcpy_h2d(self.ptr, a.ctypes.data, self.nbytes), "memcpy_h2d")

    def download(self, out=None):
        if out is None:
            out = np.empty(self.size, dtype=self.dtype)
        _check(_lib.spmvhip_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes), "memcpy_d2h")
        return out

    def fill_bytes(self, value):
        _check(_lib.spmvhip_memset(self.ptr, value, self.nbytes), "memset")

    def free(self):
        if self.ptr:
            _lib.spmvhip_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def make_params(kind, **kw):
    p = Params()
    p.kind = KINDS[kind] if isinstance(kind, str) else int(kind)
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError("unknown parameter %s" % k)
        setattr(p, k, v)
    return p


class HipFormat:
    """One converted matrix on the device (a Matrix_Format subclass instance)."""

    def __init__(self, handle, params):
        self._h = handle
        self.params = params
        info = self.info()
        self.format_name = info["format_name"]
        self.m, self.n, self.nnz = info["m"], info["n"], info["nnz"]
        self.mem_footprint = info["mem_footprint"]
        self.csr_mem_footprint = info["csr_mem_footprint"]

    # -- Matrix_Format ---------------------------------------------------
    def spmv(self, x, y):
        """Host x (n) and y (m, written in place) as the reference driver passes them."""
        assert x.dtype == np.float64 and y.dtype == np.float64
        assert x.flags.c_contiguous and y.flags.c_contiguous
        assert x.size >= self.n and y.size >= self.m
        _check(_lib.spmvhip_spmv(self._h, x.ctypes.data, y.ctypes.data), "spmvhip_spmv")

    def statistics_start(self):
        _check(_lib.spmvhip_statistics_start(self._h), "statistics_start")

    def statistics_print_data(self):
        buf = ctypes.create_string_buffer(512)
        _lib.spmvhip_statistics_print_data(self._h, buf, 512)
        return buf.value.decode()

    # -- device-resident path ---------------------------------------------
    def spmv_device(self, dx, dy, stream=None):
        _check(_lib.spmvhip_spmv_device(self._h, dx.ptr, dy.ptr, stream), "spmvhip_spmv_device")

    def time(self, dx, dy, warmup, iters):
        ms = (ctypes.c_float * max(iters, 1))()
        tot = ctypes.c_float()
        _check(_lib.spmvhip_time_spmv(self._h, dx.ptr, dy.ptr, warmup, iters, ms, ctypes.byref(tot)),
               "spmvhip_time_spmv")
        return np.array(ms[:iters], dtype=np.float64)

    def info(self):
        i = Info()
        _check(_lib.spmvhip_get_info(self._h, ctypes.byref(i)), "get_info")
        d = {f[0]: getattr(i, f[0]) for f in Info._fields_}
        d["format_name"] = d["format_name"].decode()
        return d

    def export(self, which, dtype):
        n = ctypes.c_size_t(0)
        _check(_lib.spmvhip_export(self._h, which, None, ctypes.byref(n)), "export")
        out = np.empty(n.value // np.dtype(dtype).itemsize, dtype=dtype)
        n2 = ctypes.c_size_t(out.nbytes)
        _check(_lib.spmvhip_export(self._h, which, out.ctypes.data, ctypes.byref(n2)), "export")
        return out

    def destroy(self):
        if self._h:
            _lib.spmvhip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def csr_to_format(row_ptr, col_ind, values, m, n, nnz, symmetric=0, symmetry_expanded=0,
                  kind="csr_vec", **params):
    """csr_to_format (src/spmv_kernels/spmv_kernel.h:28) for the chosen HIP format."""
    rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
    ci = np.ascontiguousarray(col_ind, dtype=np.int32)
    va = np.ascontiguousarray(values, dtype=np.float64)
    p = make_params(kind, **params)
    h = _vp()
    rc = _lib.spmvhip_csr_to_format(rp.ctypes.data, ci.ctypes.data, va.ctypes.data, m, n, nnz,
                                    symmetric, symmetry_expanded, ctypes.byref(p), ctypes.byref(h))
    _check(rc, "csr_to_format(%s)" % kind)
    return HipFormat(h.value, p)


def csr_to_format_device(d_row_ptr, d_col_ind, d_values, m, n, nnz, kind="csr_vec", **params):
    p = make_params(kind, **params)
    h = _vp()
    rc = _lib.spmvhip_csr_to_format_device(d_row_ptr.ptr, d_col_ind.ptr, d_values.ptr, m, n, nnz,
                                           ctypes.byref(p), ctypes.byref(h))
    _check(rc, "csr_to_format_device(%s)" % kind)
    return HipFormat(h.value, p)


class Batch:
    """Many converted matrices computed by one persistent, work-queue
    scheduled launch (optionally a replayed hipGraph). The member formats and
    their device vectors must stay alive while the batch is used."""

    def __init__(self, tile_units=0):
        h = _vp()
        _check(_lib.spmvhip_batch_create(tile_units, ctypes.byref(h)), "batch_create")
        self._h = h.value
        self._keep = []

    def set_queues(self, nqueues):
        """Ticket counters the units are dealt over (a power of two, default 256)."""
        _check(_lib.spmvhip_batch_set_queues(self._h, nqueues), "batch_set_queues")

    def set_grid(self, workgroups_per_cu):
        """Resident workgroups per CU of the persistent launch (0 = default)."""
        _check(_lib.spmvhip_batch_set_grid(self._h, workgroups_per_cu), "batch_set_grid")

    def set_solo_nnz(self, nnz):
        """Members of at least nnz nonzeros run as their own kernels in the
        same launch (0: none). Before any add."""
        _check(_lib.spmvhip_batch_set_solo_nnz(self._h, int(nnz)), "batch_set_solo_nnz")

    def set_xcd_ranges(self, on):
        """One contiguous, nnz-balanced range of units per XCD: True / False,
        or -1 for the automatic choice (the default)."""
        _check(_lib.spmvhip_batch_set_xcd_ranges(self._h, -1 if on == -1 else int(bool(on))),
               "batch_set_xcd_ranges")

    def add(self, fmt, dx, dy):
        _check(_lib.spmvhip_batch_add(self._h, fmt._h, dx.ptr, dy.ptr), "batch_add")
        self._keep.append((fmt, dx, dy))

    def finalize(self, use_graph=True):
        _check(_lib.spmvhip_batch_finalize(self._h, 1 if use_graph else 0), "batch_finalize")

    def run(self):
        _check(_lib.spmvhip_batch_run(self._h), "batch_run")

    def synchronize(self):
        _check(_lib.spmvhip_batch_synchronize(self._h), "batch_synchronize")

    def time(self, warmup, iters, per_matrix=False):
        ms = (ctypes.c_float * iters)()
        _check(_lib.spmvhip_batch_time(self._h, 1 if per_matrix else 0, warmup, iters, ms), "batch_time")
        return np.array(ms[:], dtype=np.float64)

    def info(self):
        a, b, c, d = ctypes.c_long(), ctypes.c_long(), ctypes.c_int(), ctypes.c_long()
        _check(_lib.spmvhip_batch_info(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                       ctypes.byref(d)), "batch_info")
        return {"matrices": a.value, "items": b.value, "grid": c.value, "nnz": d.value}

    def destroy(self):
        if self._h:
            _lib.spmvhip_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def gather_probe(mode, buf, n, iters, sink, clock_ghz=2.4):
    """Cycles per CU of one 64-lane load of the given pattern (see
    spmvhip_gather_probe) over the first n doubles of the device array buf."""
    out = ctypes.c_double()
    _check(_lib.spmvhip_gather_probe(mode, buf.ptr, n, iters, clock_ghz, sink.ptr, ctypes.byref(out)),
           "gather_probe")
    return out.value


def bandwidth_probe(mode, src, dst, iters):
    """Times of `iters` stream reads (mode 0) or copies (mode 1) of src."""
    ms = (ctypes.c_float * iters)()
    _check(_lib.spmvhip_bandwidth_probe(mode, src.ptr, dst.ptr if dst is not None else None, src.size,
                                        iters, ms), "bandwidth_probe")
    return np.array(ms[:], dtype=np.float64)


def csr_stats_device(d_row_ptr, d_col_ind, m, n):
    """Device structure statistics: mean / std / max row length, mean
    bandwidth / n, neighbours per nnz, non-empty rows."""
    out = np.empty(6, dtype=np.float64)
    _check(_lib.spmvhip_csr_stats_device(d_row_ptr.ptr, d_col_ind.ptr, m, n, out.ctypes.data), "csr_stats")
    keys = ("avg_nnz_per_row", "std_nnz_per_row", "max_nnz_per_row", "avg_bw_scaled",
            "avg_num_neighbours", "nonempty_rows")
    return dict(zip(keys, out.tolist()))


def csr_stats_device_noalloc(d_row_ptr, d_col_ind, m, n, d_out6, stream=None):
    """The allocation-free variant: the six statistics into the device
    array d_out6 (6 doubles), asynchronously (synchronize before reading)."""
    _check(_lib.spmvhip_csr_stats_device_noalloc(d_row_ptr.ptr, d_col_ind.ptr, m, n, d_out6.ptr, stream),
           "csr_stats_noalloc")


class MLP:
    """FP32 MLP with hand-written HIP forward / backward / Adam kernels."""

    def __init__(self, n_in, h1, h2, n_out, seed=0):
        h = _vp()
        _check(_lib.spmvhip_mlp_create(n_in, h1, h2, n_out, seed, ctypes.byref(h)), "mlp_create")
        self._h = h.value
        self.dims = (n_in, h1, h2, n_out)
        self.nparams = _lib.spmvhip_mlp_num_params(self._h)

    def get_params(self):
        p = np.empty(self.nparams, dtype=np.float32)
        _check(_lib.spmvhip_mlp_get_params(self._h, p.ctypes.data, self.nparams), "mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(_lib.spmvhip_mlp_set_params(self._h, p.ctypes.data, p.size), "mlp_set_params")

    def predict(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.empty((X.shape[0], self.dims[3]), dtype=np.float32)
        _check(_lib.spmvhip_mlp_predict(self._h, X.ctypes.data, X.shape[0], Y.ctypes.data), "mlp_predict")
        return Y

    def gradients(self, X, T, M=None):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        Mp = None if M is None else np.ascontiguousarray(M, dtype=np.float32)
        g = np.empty(self.nparams, dtype=np.float32)
        loss = ctypes.c_float()
        _check(_lib.spmvhip_mlp_gradients(self._h, X.ctypes.data, T.ctypes.data,
                                          None if Mp is None else Mp.ctypes.data, X.shape[0],
                                          g.ctypes.data, ctypes.byref(loss)), "mlp_gradients")
        return g, loss.value

    def train(self, X, T, M=None, batch=128, epochs=1, lr=1e-3):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        Mp = None if M is None else np.ascontiguousarray(M, dtype=np.float32)
        losses = np.empty(max(epochs, 1), dtype=np.float32)
        _check(_lib.spmvhip_mlp_train(self._h, X.ctypes.data, T.ctypes.data,
                                      None if Mp is None else Mp.ctypes.data, X.shape[0], batch, epochs,
                                      lr, losses.ctypes.data), "mlp_train")
        return losses[:epochs]

    def adam_step(self, grads, lr):
        g = np.ascontiguousarray(grads, dtype=np.float32)
        _check(_lib.spmvhip_mlp_adam_step(self._h, g.ctypes.data, lr), "mlp_adam_step")

    def get_adam_state(self):
        """(m1, m2, step): Adam's moment estimates and step counter."""
        n = self.nparams
        m1, m2, step = np.empty(n, np.float32), np.empty(n, np.float32), ctypes.c_long()
        _check(_lib.spmvhip_mlp_get_adam_state(self._h, m1.ctypes.data, m2.ctypes.data, n, ctypes.byref(step)),
               "mlp_get_adam_state")
        return m1, m2, step.value

    def set_adam_state(self, m1, m2, step):
        a = np.ascontiguousarray(m1, dtype=np.float32)
        b = np.ascontiguousarray(m2, dtype=np.float32)
        _check(_lib.spmvhip_mlp_set_adam_state(self._h, a.ctypes.data, b.ctypes.data, a.size, int(step)),
               "mlp_set_adam_state")

    def destroy(self):
        if self._h:
            _lib.spmvhip_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def div_encode(row_ptr, col_ind, values, n, T=1, L=64, P=16384):
    """Host DIV-RF encoding in the reference's packet layout (no device
    needed): returns (bytes, packet byte offsets, first packet of each group)."""
    rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
    ci = np.ascontiguousarray(col_ind, dtype=np.int32)
    va = np.ascontiguousarray(values, dtype=np.float64)
    m, nnz = rp.size - 1, ci.size
    nb, npk, ng = ctypes.c_size_t(), ctypes.c_long(), ctypes.c_long()
    _check(_lib.spmvhip_div_encode(rp.ctypes.data, ci.ctypes.data, va.ctypes.data, m, n, nnz, T, L, P,
                                   None, ctypes.byref(nb), None, ctypes.byref(npk), None, ctypes.byref(ng)),
           "div_encode")
    b = np.empty(max(nb.value, 1), dtype=np.uint8)
    off = np.empty(npk.value + 1, dtype=np.int64)
    grp = np.empty(ng.value + 1, dtype=np.int32)
    _check(_lib.spmvhip_div_encode(rp.ctypes.data, ci.ctypes.data, va.ctypes.data, m, n, nnz, T, L, P,
                                   b.ctypes.data, ctypes.byref(nb), off.ctypes.data, ctypes.byref(npk),
                                   grp.ctypes.data, ctypes.byref(ng)), "div_encode")
    return b[:nb.value], off, grp


def div_group_order(row_ptr, col_ind, values, n, T=1, L=64, P=16384):
    """The packet-group order of the reference's packet reordering for the
    stream div_encode returns (no device needed)."""
    rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
    ci = np.ascontiguousarray(col_ind, dtype=np.int32)
    va = np.ascontiguousarray(values, dtype=np.float64)
    m, nnz = rp.size - 1, ci.size
    ng = ctypes.c_long()
    _check(_lib.spmvhip_div_group_order(rp.ctypes.data, ci.ctypes.data, va.ctypes.data, m, n, nnz, T, L, P, None,
                                        ctypes.byref(ng)), "div_group_order")
    order = np.empty(max(ng.value, 1), dtype=np.int32)
    _check(_lib.spmvhip_div_group_order(rp.ctypes.data, ci.ctypes.data, va.ctypes.data, m, n, nnz, T, L, P,
                                        order.ctypes.data, ctypes.byref(ng)), "div_group_order")
    return order[:ng.value]


def div_reduce_precision(values, tolerance):
    """DIV_RF's lossy value rounding (no device needed)."""
    v = np.ascontiguousarray(values, dtype=np.float64)
    out = np.empty_like(v)
    _check(_lib.spmvhip_div_reduce_precision(v.ctypes.data, v.size, float(tolerance), out.ctypes.data),
           "div_reduce_precision")
    return out


def statistics_print_labels():
    buf = ctypes.create_string_buffer(512)
    _lib.spmvhip_statistics_print_labels(buf, 512)
    return buf.value.decode()


def div_device_encode(row_ptr, col_ind, values, n, T=1, L=64, P=16384, layout=2):
    """A device layout of the DIV-RF stream (layout: 1 reference bytes, 2 x
    slots, 3 interleaved lane streams, 4 both; no device needed): returns
    (bytes, packet byte offsets)."""
    rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
    ci = np.ascontiguousarray(col_ind, dtype=np.int32)
    va = np.ascontiguousarray(values, dtype=np.float64)
    m, nnz = rp.size - 1, ci.size
    nb, npk = ctypes.c_size_t(), ctypes.c_long()
    _check(_lib.spmvhip_div_device_encode(rp.ctypes.data, ci.ctypes.data, va.ctypes.data, m, n, nnz, T, L, P,
                                          layout, None, ctypes.byref(nb), None, ctypes.byref(npk)),
           "div_device_encode")
    b = np.empty(max(nb.value, 1), dtype=np.uint8)
    off = np.empty(npk.value + 1, dtype=np.int64)
    _check(_lib.spmvhip_div_device_encode(rp.ctypes.data, ci.ctypes.data, va.ctypes.data, m, n, nnz, T, L, P,
                                          layout, b.ctypes.data, ctypes.byref(nb), off.ctypes.data,
                                          ctypes.byref(npk)), "div_device_encode")
    return b[:nb.value], off
