"""montmul_amd: batched Montgomery squaring for many odd moduli on MI355X (gfx950).

A thin Python layer over the C ABI of libmontmul_hip.so (include/montmul_hip.h). It mirrors
the reference's host API (montmul/src/batch.rs, montmul/src/transform.rs) with Python
integers in place of `rug::Integer`:

    ModBatch(backend, moduli, base)      ModBatch::new          (batch.rs:62-105)
      .load / .load_raw / .set_exponent  (batch.rs:125-149)
      .run / .remaining / .mul           (batch.rs:152-166)
      .read_raw / .read / .is_one        (batch.rs:168-183)
    fermat(backend, moduli, base, chunk) (batch.rs:187-198)
    Params.for_bits / Params.check_exact (batch.rs:20-31)
    Hip(...)                             the GPU backend (replaces `Gpu`, gpu/mod.rs)

Errors that the reference reports with `Err` or a panic raise MontmulError; a carry out of a
product that must fit raises CarryError naming the slot ("... slot i", gpu/mod.rs:179-184).
The library must be built (python -m montmul_amd.build); there is no fallback path.
"""

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("MONTMUL_LIB") or os.path.join(_HERE, "libmontmul_hip.so")

__all__ = ["MontmulError", "CarryError", "Backend", "Hip", "Multi", "ModBatch", "Params", "fermat",
           "score_workload", "lib", "LIB_PATH", "MM_OK", "MM_EFAULT"]

MM_OK = -1
MM_EFAULT = -2


class MontmulError(RuntimeError):
    pass


class CarryError(MontmulError):
    """A carry out of a product that must fit; `slot` names the first such slot and `slots`
    every slot the failing run (or mul) hit."""

    def __init__(self, msg, slot, slots=None):
        super().__init__(msg)
        self.slot = slot
        self.slots = list(slots) if slots else [slot]


class BatchSetup(C.Structure):
    _fields_ = [("k", C.c_uint32), ("len", C.c_uint32), ("base", C.c_uint32), ("slots", C.c_uint32),
                ("n", C.POINTER(C.c_uint32)), ("nprime", C.POINTER(C.c_uint32)),
                ("two_n", C.POINTER(C.c_uint32))]


_CREATE = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(BatchSetup), C.POINTER(C.c_void_p))
_DESTROY = C.CFUNCTYPE(None, C.c_void_p)
_LOAD = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_size_t, C.POINTER(C.c_uint32))
_SETEXP = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_size_t, C.POINTER(C.c_uint32), C.c_size_t, C.c_uint64)
_RUN = C.CFUNCTYPE(C.c_int64, C.c_void_p, C.c_uint64)
_MUL = C.CFUNCTYPE(C.c_int64, C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t)
_READ = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_size_t, C.POINTER(C.c_uint32))
_REMAINING = C.CFUNCTYPE(C.c_uint64, C.c_void_p, C.c_size_t)
_SETMOD = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_size_t, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                      C.POINTER(C.c_uint32))
_RUNGROUP = C.CFUNCTYPE(C.c_int64, C.POINTER(C.c_void_p), C.c_size_t, C.c_uint64, C.POINTER(C.c_size_t))
_FAILED = C.CFUNCTYPE(C.c_size_t, C.c_void_p, C.POINTER(C.c_uint32), C.c_size_t)
_LASTERR = C.CFUNCTYPE(C.c_char_p)
_SETONES = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_size_t, C.c_size_t, C.POINTER(C.c_uint32))
_ISONEMANY = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_uint32), C.c_size_t, C.POINTER(C.c_int32))


class BackendTable(C.Structure):
    """`mm_backend`: the backend as a table of functions (transform.rs:11-20, 61-77)."""
    _fields_ = [("self", C.c_void_p), ("create", _CREATE), ("destroy", _DESTROY), ("load", _LOAD),
                ("set_exponent", _SETEXP), ("run", _RUN), ("mul", _MUL), ("read", _READ),
                ("remaining", _REMAINING), ("set_modulus", _SETMOD), ("run_group", _RUNGROUP),
                ("failed_slots", _FAILED), ("last_error", _LASTERR),
                ("bound_lo", C.c_uint64), ("bound_hi", C.c_uint64),
                ("min_len", C.c_uint32), ("max_len", C.c_uint32), ("full_k", C.c_int32),
                ("set_ones", _SETONES), ("is_one_many", _ISONEMANY)]


_OPEN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(BackendTable))
_CLOSE = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(BackendTable))


class Options(C.Structure):
    _fields_ = [("device", C.c_int32), ("force_global", C.c_int32), ("wraparound", C.c_int32),
                ("max_launch_ms", C.c_uint32), ("prefer_lds", C.c_int32), ("autotune", C.c_int32),
                ("tile", C.c_uint32), ("slots_per_cu", C.c_uint32)]


class ServerOptions(C.Structure):
    """`mm_server_options`."""
    _fields_ = [("base", C.c_uint32), ("pool_slots", C.c_uint32), ("target_ms", C.c_double),
                ("max_chunk", C.c_uint32), ("keep_raw", C.c_int32), ("hbm_bytes", C.c_uint64)]


class ServerResult(C.Structure):
    """`mm_server_result`: one completed request."""
    _fields_ = [("id", C.c_int64), ("prp", C.c_int32), ("len", C.c_uint32), ("bits", C.c_uint64),
                ("submit_s", C.c_double), ("admit_s", C.c_double), ("done_s", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("admitted", C.c_uint64), ("completed", C.c_uint64),
                ("failed", C.c_uint64), ("rounds", C.c_uint64), ("squarings", C.c_uint64),
                ("waiting", C.c_uint64), ("busy_ms", C.c_double), ("pools", C.c_uint32),
                ("hbm_used", C.c_uint64), ("releases", C.c_uint64)]


class PlanInfo(C.Structure):
    _fields_ = [("global_", C.c_int32), ("threads", C.c_uint32), ("lds_bytes", C.c_uint32),
                ("slots_per_wg", C.c_uint32), ("kernel", C.c_uint32), ("tuned", C.c_int32),
                ("tile", C.c_uint32), ("slots_per_cu", C.c_uint32)]


def _load_lib():
    if not os.path.exists(LIB_PATH):
        raise MontmulError(f"{LIB_PATH} is not built; run `python -m montmul_amd.build`")
    L = C.CDLL(LIB_PATH)
    u32p, u64p, vp = C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.c_void_p
    sig = {
        "mm_device_count": (C.c_int, []),
        "mm_hip_open": (C.c_int, [C.POINTER(Options), C.POINTER(BackendTable)]),
        "mm_hip_close": (None, [C.POINTER(BackendTable)]),
        "mm_last_error": (C.c_char_p, []),
        "mm_host_last_error": (C.c_char_p, []),
        "mm_run_bare": (C.c_int64, [vp, C.c_uint64]),
        "mm_push": (C.c_int, [vp]),
        "mm_plan": (C.c_int, [vp, C.POINTER(PlanInfo)]),
        "mm_free_memory": (C.c_uint64, [vp]),
        "mm_cu_count": (C.c_int, [vp]),
        "mm_max_lds": (C.c_uint64, [vp]),
        "mm_device_name": (C.c_int, [vp, C.c_char_p, C.c_size_t]),
        "mm_last_run_kernel_ms": (C.c_int, [vp, C.POINTER(C.c_double), C.POINTER(C.c_uint32)]),
        "mm_mul_noalloc": (C.c_int64, [vp, C.c_size_t, C.c_size_t, C.c_size_t]),
        "mm_multi_open": (C.c_int, [C.POINTER(BackendTable), C.c_size_t, C.POINTER(BackendTable)]),
        "mm_multi_close": (None, [C.POINTER(BackendTable)]),
        "mm_multi_parts": (C.c_size_t, [C.POINTER(BackendTable)]),
        "mm_multi_last_error": (C.c_char_p, []),
        "mm_probe_issue": (C.c_int, [C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.POINTER(C.c_double), u64p]),
        "mm_probe_last_error": (C.c_char_p, []),
        "mm_read_product": (C.c_int, [vp, C.c_size_t, u32p]),
        "mm_params_for_bits": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, u32p, u32p]),
        "mm_params_check_exact": (C.c_int, [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64]),
        "mm_modbatch_new": (C.c_int, [C.POINTER(BackendTable), C.c_size_t, u32p, C.POINTER(C.c_size_t),
                                      C.POINTER(C.c_int8), C.c_uint32, C.POINTER(vp)]),
        "mm_modbatch_free": (None, [vp]),
        "mm_modbatch_params": (None, [vp, u32p, u32p]),
        "mm_modbatch_slots": (C.c_size_t, [vp]),
        "mm_modbatch_state": (vp, [vp]),
        "mm_modbatch_load": (C.c_int, [vp, C.c_size_t, C.c_int, u32p, C.c_size_t]),
        "mm_modbatch_load_raw": (C.c_int, [vp, C.c_size_t, C.c_int, u32p, C.c_size_t]),
        "mm_modbatch_set_exponent": (C.c_int, [vp, C.c_size_t, C.c_int, u32p, C.c_size_t]),
        "mm_modbatch_run": (C.c_int64, [vp, C.c_uint64]),
        "mm_modbatch_remaining": (C.c_uint64, [vp, C.c_size_t]),
        "mm_modbatch_mul": (C.c_int64, [vp, C.c_size_t, C.c_size_t, C.c_size_t]),
        "mm_modbatch_failed_slots": (C.c_size_t, [vp, u32p, C.c_size_t]),
        "mm_modbatch_check_pow_range": (C.c_int64, [vp, C.c_uint32, C.c_size_t, C.c_size_t, u32p,
                                                     C.POINTER(C.c_size_t), C.c_uint32]),
        "mm_mixbatch_failed_slots": (C.c_size_t, [vp, u32p, C.c_size_t]),
        "mm_failed_slots": (C.c_size_t, [vp, u32p, C.c_size_t]),
        "mm_modbatch_read_raw": (C.c_int, [vp, C.c_size_t, u32p]),
        "mm_modbatch_read": (C.c_int, [vp, C.c_size_t, u32p]),
        "mm_modbatch_is_one": (C.c_int, [vp, C.c_size_t]),
        "mm_modbatch_is_one_many": (C.c_int, [vp, C.POINTER(C.c_uint32), C.c_size_t, C.POINTER(C.c_int32)]),
        "mm_modbatch_modulus": (C.c_int, [vp, C.c_size_t, u32p]),
        "mm_modbatch_set_modulus": (C.c_int, [vp, C.c_size_t, C.c_int, u32p, C.c_size_t]),
        "mm_set_modulus": (C.c_int, [vp, C.c_size_t, u32p, u32p, u32p]),
        "mm_server_new": (C.c_int, [C.POINTER(ServerOptions), _OPEN, _CLOSE, vp, C.POINTER(vp)]),
        "mm_server_new_hip": (C.c_int, [C.POINTER(ServerOptions), C.POINTER(Options), C.POINTER(vp)]),
        "mm_server_submit": (C.c_int64, [vp, u32p, C.c_size_t]),
        "mm_server_submit_exp": (C.c_int64, [vp, u32p, C.c_size_t, u32p, C.c_size_t]),
        "mm_server_poll": (C.c_size_t, [vp, C.POINTER(ServerResult), C.c_size_t, C.c_double]),
        "mm_server_raw": (C.c_int, [vp, C.c_int64, u32p, C.c_size_t, C.POINTER(C.c_size_t)]),
        "mm_server_stats_get": (C.c_int, [vp, C.POINTER(ServerStats)]),
        "mm_server_last_error": (C.c_char_p, [vp]),
        "mm_server_free": (None, [vp]),
        "mm_fermat": (C.c_int, [C.POINTER(BackendTable), C.c_size_t, u32p, C.POINTER(C.c_size_t),
                                C.POINTER(C.c_int8), C.c_uint32, C.c_uint64, C.POINTER(C.c_uint8)]),
        "mm_score_workload": (C.c_int, [C.c_uint32, C.c_size_t, C.c_uint32, C.c_uint64, u32p, u32p]),
        "mm_modbatch_check_pow": (C.c_int, [vp, C.c_size_t, C.c_uint32, u32p, C.c_size_t]),
        "mm_server_slot_bytes": (C.c_uint64, [C.c_uint32]),
        "mm_run_group": (C.c_int64, [C.POINTER(vp), C.c_size_t, C.c_uint64, C.POINTER(C.c_size_t)]),
        "mm_mixbatch_new": (C.c_int, [C.POINTER(BackendTable), C.c_size_t, u32p, C.POINTER(C.c_size_t),
                                      C.POINTER(C.c_int8), C.c_uint32, C.POINTER(vp)]),
        "mm_mixbatch_free": (None, [vp]),
        "mm_mixbatch_slots": (C.c_size_t, [vp]),
        "mm_mixbatch_groups": (C.c_size_t, [vp]),
        "mm_mixbatch_params": (C.c_int, [vp, C.c_size_t, u32p, u32p]),
        "mm_mixbatch_load": (C.c_int, [vp, C.c_size_t, C.c_int, u32p, C.c_size_t]),
        "mm_mixbatch_load_raw": (C.c_int, [vp, C.c_size_t, C.c_int, u32p, C.c_size_t]),
        "mm_mixbatch_set_exponent": (C.c_int, [vp, C.c_size_t, C.c_int, u32p, C.c_size_t]),
        "mm_mixbatch_run": (C.c_int64, [vp, C.c_uint64]),
        "mm_mixbatch_remaining": (C.c_uint64, [vp, C.c_size_t]),
        "mm_mixbatch_read_raw": (C.c_int, [vp, C.c_size_t, u32p]),
        "mm_mixbatch_read": (C.c_int, [vp, C.c_size_t, u32p]),
        "mm_mixbatch_is_one": (C.c_int, [vp, C.c_size_t]),
        "mm_mixbatch_is_one_many": (C.c_int, [vp, C.POINTER(C.c_uint32), C.c_size_t, C.POINTER(C.c_int32)]),
        "mm_mixbatch_check_pow": (C.c_int, [vp, C.c_size_t, C.c_uint32, u32p, C.c_size_t]),
        "mm_mixbatch_group_state": (vp, [vp, C.c_size_t]),
        "mm_t_upload": (C.c_int, [vp, u32p, C.c_uint32, C.c_uint32, C.POINTER(vp)]),
        "mm_t_load_forward": (C.c_int, [vp, u32p, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(vp)]),
        "mm_t_download": (C.c_int, [vp, u64p]),
        "mm_t_download_raw": (C.c_int, [vp, u32p, u64p]),
        "mm_t_upload_scales": (C.c_int, [vp, u32p, C.c_uint32, C.POINTER(vp)]),
        "mm_t_forward": (C.c_int, [vp, C.c_uint32, C.c_uint32]),
        "mm_t_inverse": (C.c_int, [vp, C.c_uint32, C.c_uint32]),
        "mm_t_pointwise": (C.c_int, [vp, vp, vp, vp, C.c_uint32, C.c_uint32]),
        "mm_t_free": (None, [vp]),
        "mm_t_free_scales": (None, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


class _Lib:
    """The loaded library, opened on first use (so that `montmul_amd.build` can run first)."""

    _handle = None

    def __getattr__(self, name):
        if _Lib._handle is None:
            _Lib._handle = _load_lib()
        return getattr(_Lib._handle, name)


lib = _Lib()


def _host_error():
    return (lib.mm_host_last_error() or b"").decode(errors="replace")


# ---- integers <-> limbs --------------------------------------------------------------------

def to_limbs(x, k=None):
    """Little-endian 32-bit limbs of |x| (k limbs when given; at least one)."""
    mag = abs(x)
    nl = max(1, (mag.bit_length() + 31) // 32) if k is None else k
    if mag.bit_length() > 32 * nl:
        raise MontmulError(f"value does not fit {nl} limbs")
    return (C.c_uint32 * nl).from_buffer_copy(mag.to_bytes(4 * nl, "little"))


def from_limbs(buf):
    return int.from_bytes(bytes(buf), "little")


def _sign(x):
    return (x > 0) - (x < 0)


def _pack_ints(values):
    """Concatenated limbs, offsets and signs of several integers (one C call each side)."""
    chunks, offsets, signs = [], [0], []
    for v in values:
        mag = abs(int(v))
        nl = max(1, (mag.bit_length() + 31) // 32)
        chunks.append(mag.to_bytes(4 * nl, "little"))
        offsets.append(offsets[-1] + nl)
        signs.append(_sign(int(v)))
    data = b"".join(chunks)
    limbs = (C.c_uint32 * max(1, offsets[-1])).from_buffer_copy(data + b"\0" * (4 * max(1, offsets[-1]) - len(data)))
    return limbs, (C.c_size_t * len(offsets))(*offsets), (C.c_int8 * max(1, len(signs)))(*signs)


# ---- backends -----------------------------------------------------------------------------

class Backend:
    """A backend given by its function table (`mm_backend`); `table` is a BackendTable."""

    table = None

    def coefficient_bound(self):  # Limits::coefficient_bound (transform.rs:13)
        return self.table.bound_lo | (self.table.bound_hi << 64)

    def supported_lengths(self):  # Limits::supported_lengths (transform.rs:15)
        out, length = [], self.table.min_len
        while length and length <= self.table.max_len:
            out.append(length)
            length *= 2
        return out

    def full_k(self):  # Limits::full_k (transform.rs:17)
        return bool(self.table.full_k)


class Hip(Backend):
    """The MI355X backend (replaces `Gpu`/`GpuOptions`, gpu/mod.rs:28-111)."""

    def __init__(self, device=0, force_global=False, wraparound=True, max_launch_ms=0, prefer_lds=False,
                 autotune=False, tile=0, slots_per_cu=0):
        """autotune: time every plan that fits each new batch's length (16 steps on its own
        slots) and keep the fastest, cached per length and slot-count octave. Without it, `tile`
        (the register kernel's exchange tile in global mode, in elements; 0 = default) and
        `slots_per_cu` (1 = one slot per CU; 0 = as many as fit) fix those choices."""
        self.options = Options(device, int(bool(force_global)), int(bool(wraparound)), max_launch_ms,
                               int(bool(prefer_lds)), int(bool(autotune)), tile, slots_per_cu)
        self.table = BackendTable()
        if lib.mm_hip_open(C.byref(self.options), C.byref(self.table)) != 0:
            raise MontmulError((lib.mm_last_error() or b"").decode())
        self._open = True

    @staticmethod
    def device_count():
        n = lib.mm_device_count()
        if n < 0:
            raise MontmulError((lib.mm_last_error() or b"").decode())
        return n

    def close(self):
        if getattr(self, "_open", False):
            lib.mm_hip_close(C.byref(self.table))
            self._open = False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def free_memory(self):
        return lib.mm_free_memory(self.table.self)

    def cu_count(self):
        return lib.mm_cu_count(self.table.self)

    def max_lds(self):
        return lib.mm_max_lds(self.table.self)

    def device_name(self):
        buf = C.create_string_buffer(256)
        lib.mm_device_name(self.table.self, buf, 256)
        return buf.value.decode()

    # GPU-only extras on a batch's state.
    @staticmethod
    def plan_of(batch):
        p = PlanInfo()
        lib.mm_plan(batch.state, C.byref(p))
        return {"global": bool(p.global_), "threads": p.threads, "lds_bytes": p.lds_bytes,
                "slots_per_wg": p.slots_per_wg, "kernel": p.kernel, "tuned": bool(p.tuned),
                "tile": p.tile, "slots_per_cu": p.slots_per_cu}

    @staticmethod
    def run_bare(batch, bits):
        r = lib.mm_run_bare(batch.state, bits)
        _check_run(r, lambda: (lib.mm_last_error() or b"").decode())

    @staticmethod
    def read_product(batch, slot):
        """T = c x^2 of the slot's latest step (global mode), as an integer."""
        k = batch.params().k
        buf = (C.c_uint32 * (2 * k))()
        if lib.mm_read_product(batch.state, slot, buf) != 0:
            raise _hip_error()
        return from_limbs(buf)

    @staticmethod
    def push(batch):
        if lib.mm_push(batch.state) != 0:
            raise MontmulError((lib.mm_last_error() or b"").decode())

    @staticmethod
    def last_run_kernel_ms(batch):
        ms, n = C.c_double(), C.c_uint32()
        lib.mm_last_run_kernel_ms(batch.state, C.byref(ms), C.byref(n))
        return ms.value, n.value


class Multi(Backend):
    """Several backends as one (`Multi`, multi.rs:11-156): a batch's slots are split into
    contiguous shards, one per part, each created and run on its own host thread; a mul across
    shards runs on the host. Failures name "shard s, slot g" with g the batch's slot. The parts
    (for example `Hip()` contexts, on one device or one per device) stay alive with it."""

    def __init__(self, parts):
        self.parts = list(parts)
        arr = (BackendTable * max(1, len(self.parts)))()
        for i, p in enumerate(self.parts):
            arr[i] = p.table
        self._arr = arr
        self.table = BackendTable()
        if lib.mm_multi_open(arr, len(self.parts), C.byref(self.table)) != 0:
            raise MontmulError((lib.mm_multi_last_error() or b"").decode())

    def close(self):
        if self.table is not None and self.table.self:
            lib.mm_multi_close(C.byref(self.table))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _hip_error():
    return MontmulError((lib.mm_last_error() or b"").decode(errors="replace"))


class TBuf:
    """`count` transform vectors of length `len` in device memory (GpuBuf, gpu/mod.rs:228-236)."""

    def __init__(self, handle, length, count):
        self._h, self.len, self.count = handle, length, count

    def free(self):
        if self._h:
            lib.mm_t_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class TScales:
    """Per-vector scalars on the device (Transform::Scales, transform.rs:26)."""

    def __init__(self, handle, count):
        self._h, self.count = handle, count

    def __del__(self):
        try:
            if self._h:
                lib.mm_t_free_scales(self._h)
                self._h = None
        except Exception:
            pass


def _u32_array(values, count):
    """A ctypes uint32 array of `count` (at least 1) words from any sequence or numpy array."""
    try:
        import numpy as np
    except ImportError:  # plain Python fallback for the host layer
        arr = (C.c_uint32 * max(count, 1))()
        for i, v in enumerate(values):
            arr[i] = int(v) & 0xFFFFFFFF
        return arr
    a = np.zeros(max(count, 1), dtype=np.uint32)
    if count:
        v = values if isinstance(values, np.ndarray) else np.array([int(x) & 0xFFFFFFFF for x in values], dtype=np.uint64)
        a[:count] = v.astype(np.uint64, copy=False) & 0xFFFFFFFF
    return (C.c_uint32 * a.size).from_buffer(a)


def _transform_methods(cls):
    """The transform layer of a Hip table (trait Transform, transform.rs:22-46; Gpu's impl,
    gpu/mod.rs:286-319, with load_forward and download_raw, gpu/mod.rs:212-225)."""

    def upload(self, coeffs, length, count):
        n = length * count
        if len(coeffs) != n:
            raise MontmulError(f"upload: {len(coeffs)} coefficients for {count} vectors of length {length}")
        h = C.c_void_p()
        if lib.mm_t_upload(self.table.self, _u32_array(coeffs, n), length, count, C.byref(h)) != 0:
            raise _hip_error()
        return TBuf(h.value, length, count)

    def load_forward(self, limbs, k, length, count):
        h = C.c_void_p()
        if lib.mm_t_load_forward(self.table.self, _u32_array(limbs, k * count), k, length, count, C.byref(h)) != 0:
            raise _hip_error()
        return TBuf(h.value, length, count)

    def download_words(self, buf):
        """Exact coefficients as (low, high) 64-bit word pairs, len * count of them."""
        out = (C.c_uint64 * (2 * buf.len * buf.count))()
        if lib.mm_t_download(buf._h, out) != 0:
            raise _hip_error()
        return out

    def download(self, buf):
        w = self.download_words(buf)
        return [w[2 * i] | (w[2 * i + 1] << 64) for i in range(buf.len * buf.count)]

    def download_raw(self, buf):
        """(f1, f2): the elements without CRT as flat (re, im) arrays (Gpu::download_raw)."""
        el = buf.len // 2 * buf.count
        f1, f2 = (C.c_uint32 * (2 * el))(), (C.c_uint64 * (2 * el))()
        if lib.mm_t_download_raw(buf._h, f1, f2) != 0:
            raise _hip_error()
        return f1, f2

    def upload_scales(self, scales):
        h = C.c_void_p()
        if lib.mm_t_upload_scales(self.table.self, _u32_array(scales, len(scales)), len(scales), C.byref(h)) != 0:
            raise _hip_error()
        return TScales(h.value, len(scales))

    def forward(self, buf, length, count):
        if lib.mm_t_forward(buf._h, length, count) != 0:
            raise _hip_error()

    def inverse(self, buf, length, count):
        if lib.mm_t_inverse(buf._h, length, count) != 0:
            raise _hip_error()

    def pointwise(self, out, a, b, scales, length, count):
        s = scales._h if scales is not None else None
        if lib.mm_t_pointwise(out._h, a._h, b._h, s, length, count) != 0:
            raise _hip_error()

    for f in (upload, load_forward, download_words, download, download_raw, upload_scales, forward, inverse, pointwise):
        setattr(cls, f.__name__, f)
    return cls


_transform_methods(Hip)


def _check_run(r, msg, failed=None):
    """Raises for a run's or mul's result: CarryError naming the first failing slot (with every
    failing slot from `failed()`, when given), MontmulError for any other failure."""
    if r == MM_OK:
        return
    if r >= 0:
        text = msg()
        raise CarryError(text, r, failed() if failed else None)
    raise MontmulError(msg())


def _failed_list(fn, handle):
    cnt = fn(handle, None, 0)
    buf = (C.c_uint32 * max(1, cnt))()
    fn(handle, buf, cnt)
    return [buf[i] for i in range(cnt)]


# ---- host API -----------------------------------------------------------------------------

class Params:
    """`Params` (batch.rs:9-31)."""

    W = 32

    def __init__(self, k, length):
        self.k, self.len = k, length

    def __repr__(self):
        return f"Params(k={self.k}, len={self.len})"

    def __eq__(self, o):
        return isinstance(o, Params) and (self.k, self.len) == (o.k, o.len)

    @staticmethod
    def for_bits(bits, lengths):
        lengths = list(lengths)
        k, ln = C.c_uint32(), C.c_uint32()
        if lib.mm_params_for_bits(bits, min(lengths), max(lengths), C.byref(k), C.byref(ln)) != 0:
            raise MontmulError(_host_error())
        return Params(k.value, ln.value)

    def check_exact(self, c, bound):
        if lib.mm_params_check_exact(self.k, c, bound & (2**64 - 1), bound >> 64) != 0:
            raise MontmulError(_host_error())


class ModBatch:
    """`ModBatch<B>` (batch.rs:53-185) over any backend table."""

    def __init__(self, backend, moduli, base):
        self.backend = backend
        moduli = [int(m) for m in moduli]
        limbs, offsets, signs = _pack_ints(moduli)
        h = C.c_void_p()
        if lib.mm_modbatch_new(C.byref(backend.table), len(moduli), limbs, offsets, signs, base, C.byref(h)) != 0:
            raise MontmulError(_host_error())
        self._h = h
        self._moduli = moduli
        k, ln = C.c_uint32(), C.c_uint32()
        lib.mm_modbatch_params(h, C.byref(k), C.byref(ln))
        self._params = Params(k.value, ln.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib.mm_modbatch_free(h)
            self._h = None

    def free(self):
        self.__del__()

    @property
    def state(self):
        return lib.mm_modbatch_state(self._h)

    def params(self):
        return self._params

    def slots(self):
        return len(self._moduli)

    def modulus(self, i):
        return self._moduli[i]

    def set_modulus(self, i, n):
        """Replaces slot i's modulus by another needing this batch's L; clears its exponent
        (admission into a running batch; no reference counterpart)."""
        n = int(n)
        buf = to_limbs(abs(n))
        if lib.mm_modbatch_set_modulus(self._h, i, _sign(n), buf, len(buf)) != 0:
            raise MontmulError(_host_error())
        self._moduli[i] = n

    def load(self, i, x):
        x = int(x)
        buf = to_limbs(x)
        if lib.mm_modbatch_load(self._h, i, _sign(x), buf, len(buf)) != 0:
            raise MontmulError(_host_error())

    def load_raw(self, i, xm):
        xm = int(xm)
        buf = to_limbs(xm)
        if lib.mm_modbatch_load_raw(self._h, i, _sign(xm), buf, len(buf)) != 0:
            raise MontmulError(_host_error())

    def set_exponent(self, i, e):
        e = int(e)
        buf = to_limbs(e)
        if lib.mm_modbatch_set_exponent(self._h, i, _sign(e), buf, len(buf)) != 0:
            raise MontmulError(_host_error())

    def run(self, bits):
        _check_run(lib.mm_modbatch_run(self._h, bits), _host_error, self.failed_slots)

    def failed_slots(self):
        """Every slot whose carry check failed in the latest run or mul (no reference
        counterpart: the reference panics on the first)."""
        return _failed_list(lib.mm_modbatch_failed_slots, self._h)

    def remaining(self, i):
        return lib.mm_modbatch_remaining(self._h, i)

    def mul(self, out, a, b):
        _check_run(lib.mm_modbatch_mul(self._h, out, a, b), _host_error, self.failed_slots)

    def read_raw(self, i):
        buf = (C.c_uint32 * self._params.k)()
        if lib.mm_modbatch_read_raw(self._h, i, buf) != 0:
            raise MontmulError(_host_error())
        return from_limbs(buf)

    def read(self, i):
        buf = (C.c_uint32 * self._params.k)()
        if lib.mm_modbatch_read(self._h, i, buf) != 0:
            raise MontmulError(_host_error())
        return from_limbs(buf)

    def check_pow(self, i, base, e):
        """No bits left and read(i) == base^e mod N, checked with GMP (score.rs:107-115)."""
        buf = to_limbs(int(e))
        r = lib.mm_modbatch_check_pow(self._h, i, base, buf, len(buf))
        if r < 0:
            raise MontmulError(_host_error())
        return bool(r)

    def check_pow_range(self, base, first, exps, threads=1):
        """check_pow on slots first .. first+len(exps)-1 (GMP powm on `threads` host threads):
        None when all match, else the lowest slot that does not."""
        limbs, offsets, _ = _pack_ints([int(e) for e in exps])
        r = lib.mm_modbatch_check_pow_range(self._h, base, first, len(exps), limbs, offsets, threads)
        if r == -2:
            raise MontmulError(_host_error())
        return None if r == -1 else r

    def is_one(self, i):
        r = lib.mm_modbatch_is_one(self._h, i)
        if r < 0:
            raise MontmulError(_host_error())
        return bool(r)

    def is_one_many(self, slots):
        """is_one for many slots at once (on the device on the Hip backend): a list of bools."""
        slots = list(slots)
        arr = (C.c_uint32 * max(1, len(slots)))(*slots)
        out = (C.c_int32 * max(1, len(slots)))()
        if lib.mm_modbatch_is_one_many(self._h, arr, len(slots), out):
            raise MontmulError(_host_error())
        return [bool(out[j]) for j in range(len(slots))]


class MixBatch:
    """Moduli of different sizes in one batch: each gets the Params of its own size
    (batch.rs:20-25), the moduli of one length form one ModBatch, and run() steps every length
    together, through one persistent launch on the Hip backend (group by group on a backend
    without a joint run). Same per-slot methods as ModBatch. No reference counterpart:
    ModBatch::new rejects mixed lengths (batch.rs:64-77)."""

    def __init__(self, backend, moduli, base):
        self.backend = backend
        moduli = [int(m) for m in moduli]
        limbs, offsets, signs = _pack_ints(moduli)
        h = C.c_void_p()
        if lib.mm_mixbatch_new(C.byref(backend.table), len(moduli), limbs, offsets, signs, base, C.byref(h)) != 0:
            raise MontmulError(_host_error())
        self._h = h
        self._moduli = moduli
        self._params = []
        for i in range(len(moduli)):
            k, ln = C.c_uint32(), C.c_uint32()
            lib.mm_mixbatch_params(h, i, C.byref(k), C.byref(ln))
            self._params.append(Params(k.value, ln.value))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib.mm_mixbatch_free(h)
            self._h = None

    def free(self):
        self.__del__()

    def params(self, i):
        if not 0 <= i < len(self._params):
            raise MontmulError(f"slot {i} out of range")
        return self._params[i]

    def slots(self):
        return len(self._moduli)

    def groups(self):
        return lib.mm_mixbatch_groups(self._h)

    def group_state(self, g):
        """The backend state of length group g (ascending L)."""
        return lib.mm_mixbatch_group_state(self._h, g)

    def modulus(self, i):
        return self._moduli[i]

    def load(self, i, x):
        x = int(x)
        buf = to_limbs(x)
        if lib.mm_mixbatch_load(self._h, i, _sign(x), buf, len(buf)) != 0:
            raise MontmulError(_host_error())

    def load_raw(self, i, xm):
        xm = int(xm)
        buf = to_limbs(xm)
        if lib.mm_mixbatch_load_raw(self._h, i, _sign(xm), buf, len(buf)) != 0:
            raise MontmulError(_host_error())

    def set_exponent(self, i, e):
        e = int(e)
        buf = to_limbs(e)
        if lib.mm_mixbatch_set_exponent(self._h, i, _sign(e), buf, len(buf)) != 0:
            raise MontmulError(_host_error())

    def run(self, bits):
        _check_run(lib.mm_mixbatch_run(self._h, bits), _host_error, self.failed_slots)

    def failed_slots(self):
        """Every slot whose carry check failed in the latest run."""
        return _failed_list(lib.mm_mixbatch_failed_slots, self._h)

    def remaining(self, i):
        return lib.mm_mixbatch_remaining(self._h, i)

    def read_raw(self, i):
        buf = (C.c_uint32 * self.params(i).k)()
        if lib.mm_mixbatch_read_raw(self._h, i, buf) != 0:
            raise MontmulError(_host_error())
        return from_limbs(buf)

    def read(self, i):
        buf = (C.c_uint32 * self.params(i).k)()
        if lib.mm_mixbatch_read(self._h, i, buf) != 0:
            raise MontmulError(_host_error())
        return from_limbs(buf)

    def check_pow(self, i, base, e):
        buf = to_limbs(int(e))
        r = lib.mm_mixbatch_check_pow(self._h, i, base, buf, len(buf))
        if r < 0:
            raise MontmulError(_host_error())
        return bool(r)

    def is_one(self, i):
        r = lib.mm_mixbatch_is_one(self._h, i)
        if r < 0:
            raise MontmulError(_host_error())
        return bool(r)

    def is_one_many(self, slots):
        """is_one for many slots at once (each length's slots together, on the device on Hip)."""
        slots = list(slots)
        arr = (C.c_uint32 * max(1, len(slots)))(*slots)
        out = (C.c_int32 * max(1, len(slots)))()
        if lib.mm_mixbatch_is_one_many(self._h, arr, len(slots), out):
            raise MontmulError(_host_error())
        return [bool(out[j]) for j in range(len(slots))]


def fermat(backend, moduli, base, chunk):
    """Base-`base` Fermat PRP test of every modulus (batch.rs:187-198)."""
    moduli = [int(m) for m in moduli]
    limbs, offsets, signs = _pack_ints(moduli)
    out = (C.c_uint8 * max(1, len(moduli)))()
    if lib.mm_fermat(C.byref(backend.table), len(moduli), limbs, offsets, signs, base, chunk, out) != 0:
        raise MontmulError(_host_error())
    return [bool(out[i]) for i in range(len(moduli))]


def score_workload(bits, slots, ebits, seed):
    """The score workload (score.rs:45-61): moduli and exponents from GMP's seeded generator."""
    nl, el = (bits + 31) // 32, (ebits + 31) // 32
    mod = (C.c_uint32 * (slots * nl))()
    exp = (C.c_uint32 * (slots * el))()
    if lib.mm_score_workload(bits, slots, ebits, seed, mod, exp) != 0:
        raise MontmulError(_host_error())
    mb, eb = bytes(mod), bytes(exp)
    moduli = [int.from_bytes(mb[4 * nl * i:4 * nl * (i + 1)], "little") for i in range(slots)]
    exps = [int.from_bytes(eb[4 * el * i:4 * el * (i + 1)], "little") for i in range(slots)]
    return moduli, exps


class Server:
    """Base-c Fermat PRP requests served over length pools (mm_server; no reference
    counterpart). By default every pool is a HIP backend on `device`; `backend_factory` (a
    callable returning a backend such as the oracle's) replaces it, one backend per pool."""

    def __init__(self, base=3, pool_slots=256, target_ms=20.0, max_chunk=0, keep_raw=False, device=0,
                 force_global=False, backend_factory=None, hbm_bytes=0):
        """hbm_bytes: device memory all pools together may take (0: no limit), counted with
        slot_bytes(L) per slot; a length that finds no room waits for an idle pool's memory."""
        self.options = ServerOptions(base, pool_slots, target_ms, max_chunk, int(bool(keep_raw)), hbm_bytes)
        self._h = C.c_void_p()
        self._keep = []
        if backend_factory is None:
            hip = Options(device, int(bool(force_global)), 1, 0, 0, 0)
            r = lib.mm_server_new_hip(C.byref(self.options), C.byref(hip), C.byref(self._h))
        else:
            def _open(user, out):
                try:
                    be = backend_factory()
                except Exception:
                    return 1
                self._keep.append(be)
                out[0] = be.table
                return 0

            def _close(user, be):
                return None

            self._cb = (_OPEN(_open), _CLOSE(_close))
            r = lib.mm_server_new(C.byref(self.options), self._cb[0], self._cb[1], None, C.byref(self._h))
        if r != 0:
            raise MontmulError(_host_error() or (lib.mm_server_last_error(None) or b"").decode())

    @staticmethod
    def slot_bytes(length):
        """Device memory of one pool slot at transform length L in the server's accounting."""
        return lib.mm_server_slot_bytes(length)

    def submit(self, n, e=None):
        """Queues a Fermat test of N (base^(N-1) mod N), or base^e mod N for a given e >= 0.
        Returns the request's id."""
        buf = to_limbs(int(n))
        if e is None:
            i = lib.mm_server_submit(self._h, buf, len(buf))
        else:
            ebuf = to_limbs(int(e))
            i = lib.mm_server_submit_exp(self._h, buf, len(buf), ebuf, len(ebuf))
        if i < 0:
            raise MontmulError((lib.mm_server_last_error(self._h) or b"").decode())
        return i

    def poll(self, max_results=256, wait=0.0):
        """Completed requests: dicts with id, prp (1, 0, or -1 on failure), len, bits and the
        submit/admit/done times in seconds since the server started."""
        out = (ServerResult * max_results)()
        n = lib.mm_server_poll(self._h, out, max_results, wait)
        return [{"id": r.id, "prp": r.prp, "len": r.len, "bits": r.bits, "submit_s": r.submit_s,
                 "admit_s": r.admit_s, "done_s": r.done_s} for r in out[:n]]

    def raw(self, request_id, max_limbs=1 << 20):
        out = (C.c_uint32 * max_limbs)()
        nl = C.c_size_t()
        if lib.mm_server_raw(self._h, request_id, out, max_limbs, C.byref(nl)) != 0:
            raise MontmulError((lib.mm_server_last_error(self._h) or b"").decode())
        return int.from_bytes(C.string_at(out, 4 * nl.value), "little")

    def stats(self):
        st = ServerStats()
        lib.mm_server_stats_get(self._h, C.byref(st))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def close(self):
        if self._h:
            lib.mm_server_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

