"""
vulkpy Core Module on MI355X (:mod:`vulkpy_amd.vkarray`)
========================================================

* GPU
* Array, U32Array, Shape
* zeros

Same public API as vulkpy/vkarray.py (GPU 68-136, _GPUArray 139-188,
U32Array 191-256, Shape 259-278, Array 281-1521, zeros 1524-1542).

Mechanism differences:

* Data lives in HBM (``hipMalloc`` through a stream-ordered caching
  allocator), not in host-visible Vulkan memory.  ``Array.array`` is a host
  mirror that is downloaded lazily when read; ``arr[key] = value`` waits,
  writes the mirror and uploads it; after writing ``arr.array`` directly call
  ``arr.flush()`` (the reference's own flush, vkarray.py:163-167).
* Every operation enqueues kernels on the GPU's single in-order HIP stream and
  returns at once; ``job`` marks the point of the stream after them (an event
  is recorded there only if someone waits for it).  Dependencies are ordered
  by the stream instead of host waits (_vkarray.cc:430-432).
"""
from __future__ import annotations

import ctypes
import functools
import logging
import os
import weakref
from typing import Iterable, List, Optional, Tuple, Union

import numpy as np

from . import _lib
from ._lib import BINOP, REDOP, UNOP, GEMM_MODE, call, shape_array
from .vktyping import KeyType, ValueType, Resource

__all__ = ["GPU", "U32Array", "Shape", "Array", "zeros", "Job"]

logger = logging.getLogger("vulkpy")

_UINT64_MAX = (1 << 64) - 1


class _Context:
    """Process-wide HIP context of one device: stream + allocator cache."""

    def __init__(self, idx: int, priority: float):
        out = ctypes.c_void_p()
        call("vkh_gpu_create", int(idx), float(priority), ctypes.byref(out))
        self.handle = out.value
        self.idx = idx
        cus, lds, wave = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        call("vkh_gpu_info", self.handle, ctypes.byref(cus), ctypes.byref(lds), ctypes.byref(wave))
        self.num_cus, self.lds_bytes, self.wave_size = cus.value, lds.value, wave.value
        # Jobs are tickets: job k is the k-th piece of work enqueued on the
        # stream.  A HIP event is recorded only when a job is waited for or
        # queried and no recorded event covers it yet; on the in-order stream
        # an event recorded later covers every earlier ticket.
        self.ticket = 0
        self.done_ticket = 0
        self.latest: Optional[_Event] = None

    def issue(self) -> int:
        self.ticket += 1
        return self.ticket

    def covering(self, ticket: int) -> "_Event":
        """A recorded event that completes after job `ticket`."""
        e = self.latest
        if e is None or e.ticket < ticket:
            e = _Event(self)
            self.latest = e
        return e

    def mark_done(self, ticket: int) -> None:
        if ticket > self.done_ticket:
            self.done_ticket = ticket


# ---- finalizers during a graph capture -------------------------------------------
# A capturing stream cannot be waited on, and a host wait, an event's
# destruction or a pinned block's release in the capturing thread would end the
# capture with an error.  Finalizers run by garbage collection meanwhile (any
# allocation can trigger one) queue that work here instead; it runs when the
# capture ends.
_captures = 0
_deferred: List = []


def _capture_begin() -> None:
    global _captures
    _captures += 1


def _capture_end() -> None:
    global _captures
    _captures -= 1
    if _captures == 0:
        work = list(_deferred)
        del _deferred[:]
        for fn in work:
            try:
                fn()
            except Exception:
                pass


def _after_capture(fn) -> bool:
    """Queue ``fn`` for the end of the running capture; False when none runs."""
    if _captures == 0:
        return False
    _deferred.append(fn)
    return True


def _destroy_event(ev: int) -> None:
    _lib.lib.vkh_event_destroy(ev)


class _Event:
    """One recorded HIP event and the tickets it covers."""

    __slots__ = ("ev", "ticket")

    def __init__(self, ctx: "_Context"):
        ev = ctypes.c_void_p()
        call("vkh_event_record", ctx.handle, ctypes.byref(ev))
        self.ev = ev.value
        self.ticket = ctx.ticket

    def __del__(self):
        try:
            if self.ev:
                ev, self.ev = self.ev, None
                if not _after_capture(functools.partial(_destroy_event, ev)):
                    _destroy_event(ev)
        except Exception:
            pass


_contexts = {}


def _context(idx: int, priority: float) -> _Context:
    ctx = _contexts.get(idx)
    if ctx is None:
        ctx = _Context(idx, priority)
        _contexts[idx] = ctx
    return ctx


class Job:
    """
    Completion handle of enqueued GPU work (replaces vulkpy's Job,
    _vkarray.cc:392-457).

    Creating a job enqueues nothing: it is a ticket on the device's in-order
    stream, and a HIP event is recorded only when a job is waited for or
    queried (an event records a packet that costs GPU time, about 4 us, so
    one per operation would tax every step).  Such an event also covers work
    enqueued after the job, so a wait may return a little later than the
    job's own work finished; it never returns earlier.
    """

    __slots__ = ("_ctx", "_ticket")

    def __init__(self, ctx: _Context):
        self._ctx = ctx
        self._ticket = ctx.issue()

    def wait(self, timeout_ns: Optional[int] = None):
        """
        Wait for the job

        Parameters
        ----------
        timeout_ns : int, optional
            Timeout in nanoseconds.  ``RuntimeError("Timeout at Command Wait")``
            is raised when it expires (_vkarray.cc:446-456).
        """
        ctx = self._ctx
        if self._ticket <= ctx.done_ticket:
            return
        if _captures:
            raise RuntimeError("Job.wait during a graph capture")
        e = ctx.covering(self._ticket)  # held here while waiting
        t = _UINT64_MAX if timeout_ns is None else int(timeout_ns)
        _lib.check(_lib.lib.vkh_event_wait(e.ev, t), "Error at Command Wait")
        ctx.mark_done(e.ticket)

    def done(self) -> bool:
        ctx = self._ctx
        if self._ticket <= ctx.done_ticket:
            return True
        e = ctx.covering(self._ticket)
        flag = ctypes.c_int()
        call("vkh_event_query", e.ev, ctypes.byref(flag))
        if flag.value:
            ctx.mark_done(e.ticket)
        return bool(flag.value)


class GPU:
    """
    GPU instance

    Examples
    --------
    >>> import vulkpy_amd as vk
    >>> gpu = vk.GPU()

    GPU equality is checked by GPU index

    >>> gpu == vk.GPU()
    True

    Notes
    -----
    Replaces vulkpy/vkarray.py:68-136.  All ``GPU`` objects with the same index
    share one HIP stream and allocator.  ``priority`` is accepted for API
    compatibility; the stream always runs at the default priority.
    """

    def __init__(self, idx: int = 0, priority: float = 0.0):
        """
        Initialize GPU

        Parameters
        ----------
        idx : int, optional
            Index to specify one GPU from multiple GPUs. Default is ``0``.
        priority : float, optional
            GPU priority. Default is ``0.0``.

        Raises
        ------
        RuntimeError
            If the HIP library is not built or no HIP device is visible.
        """
        self._idx = idx
        self._ctx = _context(idx, priority)
        self.gpu = self._ctx
        self.canSubgroupArithmetic = True
        logger.info(f"GPU {idx}: Subgroup Arithmetic: {self.canSubgroupArithmetic}")

    def __eq__(self, other: object):
        if not isinstance(other, GPU):
            return NotImplemented
        return self._idx == other._idx

    def __hash__(self):
        return hash(self._idx)

    @property
    def _handle(self):
        return self._ctx.handle

    def _job(self) -> Job:
        return Job(self._ctx)

    def flush(self, arrays: Iterable["_GPUArray"]):
        """
        Flush buffers: upload the host mirrors of ``arrays``

        Parameters
        ----------
        arrays : iterable of Array
            Arrays to be flushed
        """
        for a in arrays:
            a._upload_host()

    def wait(self):
        """
        Wait All GPU Operations
        """
        t = self._ctx.ticket
        call("vkh_gpu_wait", self._handle)
        self._ctx.mark_done(t)

    def memory_stats(self) -> Tuple[int, int]:
        """(live, cached) bytes held by this GPU's allocator."""
        live, cached = ctypes.c_size_t(), ctypes.c_size_t()
        call("vkh_memory_stats", self._handle, ctypes.byref(live), ctypes.byref(cached))
        return live.value, cached.value

    def empty_cache(self):
        """Release cached device blocks back to the driver."""
        call("vkh_empty_cache", self._handle)


class _DeviceBuffer:
    """Device allocation of ``n`` elements of ``dtype`` (Buffer<T>, _vkarray.cc:38-130)."""

    __slots__ = ("_gpu", "ptr", "n", "dtype", "nbytes")

    def __init__(self, gpu: GPU, n: int, dtype):
        self._gpu = gpu
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        out = ctypes.c_void_p()
        call("vkh_alloc", gpu._handle, max(self.nbytes, 1), ctypes.byref(out))
        self.ptr = out.value

    def size(self) -> int:
        return self.n

    def __del__(self):
        try:
            if self.ptr:
                _lib.lib.vkh_free(self._gpu._handle, self.ptr)
                self.ptr = None
        except Exception:
            pass


def _parse_shape(shape) -> Tuple[int, ...]:
    return tuple(int(s) for s in np.asarray(shape, dtype=int).reshape(-1))


def _reshape_target(size: int, shape) -> Tuple[int, ...]:
    if isinstance(shape, (int, np.integer)):
        shape = (shape,)
    shape = [int(s) for s in shape]
    neg = [i for i, s in enumerate(shape) if s == -1]
    if len(neg) > 1:
        raise ValueError("can only specify one unknown dimension")
    if neg:
        known = int(np.prod([s for s in shape if s != -1], dtype=np.int64))
        if known == 0 or size % known != 0:
            raise ValueError(f"cannot reshape array of size {size} into shape {tuple(shape)}")
        shape[neg[0]] = size // known
    if any(s < 0 for s in shape) or int(np.prod(shape, dtype=np.int64)) != size:
        raise ValueError(f"cannot reshape array of size {size} into shape {tuple(shape)}")
    return tuple(shape)


_PIN_MIN_BYTES = 1 << 20
_UPLOAD_CHUNK_BYTES = 16 << 20


def _wait_then_drop(job, host) -> None:
    job.wait()


def _host_free(ptr: int) -> None:
    _lib.load().vkh_host_free(ptr)


def _host_release(ptr: int) -> None:
    try:
        if not _after_capture(functools.partial(_host_free, ptr)):
            _host_free(ptr)
    except Exception:
        pass


def _host_empty(n: int, dtype) -> np.ndarray:
    """
    Host mirror of ``n`` elements.  Mirrors of 1 MiB and more are page-locked
    blocks from the library's pinned pool (uploads and downloads are then
    single DMA transfers); smaller ones, or any once the pool refuses, are
    ordinary memory.  A pinned block goes back to the pool when the last view
    of it is gone.
    """
    dtype = np.dtype(dtype)
    nbytes = int(n) * dtype.itemsize
    if nbytes < _PIN_MIN_BYTES:
        return np.empty(n, dtype=dtype)
    p = ctypes.c_void_p()
    try:
        call("vkh_host_alloc", nbytes, ctypes.byref(p))
    except (MemoryError, RuntimeError):
        return np.empty(n, dtype=dtype)
    buf = (ctypes.c_char * nbytes).from_address(p.value)
    fin = weakref.finalize(buf, _host_release, p.value)
    fin.atexit = False  # the process returns page-locked memory when it ends
    return np.frombuffer(buf, dtype=dtype, count=n)


class _GPUArray(Resource):
    """Common base of Array and U32Array (vkarray.py:139-188)."""

    _dtype = np.float32

    def __init__(self, gpu: GPU):
        self._gpu: GPU = gpu

        # Pipeline job to write this Array.
        self.job: Optional[Job] = None

        # Hold temporary resources until pipeline job finish.
        self._keep: List[Resource] = []

        self._host: Optional[np.ndarray] = None
        self._host_valid = False
        self._upload_pending = False

    # ---- construction helpers -------------------------------------------------
    def _init_data(self, data, dtype):
        # np.array(data, dtype) of the reference (vkarray.py:423), copied once
        # into the host mirror and uploaded from there
        arr = np.asarray(data, dtype=dtype)
        self.shape = arr.shape
        self.buffer = _DeviceBuffer(self._gpu, arr.size, dtype)
        self._host = _host_empty(arr.size, dtype)
        chunk = _UPLOAD_CHUNK_BYTES // self._host.itemsize
        if (isinstance(self._host.base, ctypes.Array) and arr.flags.c_contiguous
                and arr.size >= 2 * chunk):
            # pinned mirror: the DMA of each chunk overlaps the host copy of
            # the next one
            src = arr.reshape(-1)
            h, itemsize = self._gpu._handle, self._host.itemsize
            for off in range(0, arr.size, chunk):
                end = min(arr.size, off + chunk)
                self._host[off:end] = src[off:end]
                call("vkh_upload", h, self._ptr + off * itemsize,
                     self._host.ctypes.data + off * itemsize, (end - off) * itemsize)
            self.job = self._gpu._job()
            self._upload_pending = True
            self._host_valid = True
            return
        np.copyto(self._host.reshape(arr.shape), arr)
        self._host_valid = True
        self._upload_host()

    def _init_shape(self, shape, dtype):
        self.shape = _parse_shape(shape)
        self.buffer = _DeviceBuffer(self._gpu, int(np.prod(self.shape, dtype=np.int64)), dtype)

    @property
    def _ptr(self) -> int:
        return self.buffer.ptr

    @property
    def size(self) -> int:
        return self.buffer.n

    # ---- synchronisation --------------------------------------------------------
    def __del__(self):
        # The device block goes back to the stream-ordered allocator without a
        # host wait; only an upload still reading the host mirror is waited for.
        try:
            if self._upload_pending and self.job is not None:
                # during a capture: wait afterwards, the mirror kept till then
                if not _after_capture(functools.partial(_wait_then_drop, self.job, self._host)):
                    self.job.wait()
        except Exception:
            pass

    def wait(self):
        """
        Wait Last Job
        """
        if self.job is not None:
            self.job.wait()
            self.job = None
        self._keep = []
        self._upload_pending = False

    def _set_job(self, keep: Iterable[Resource] = ()):
        # _upload_pending stays as it is: an upload from the host mirror may
        # still be queued, and the new ticket (later on the same in-order
        # stream) is what __del__ must wait for before the mirror's pinned
        # block can go back to the pool.
        self.job = self._gpu._job()
        self._keep = list(keep)
        self._host_valid = False

    def flush(self):
        """
        Flush Buffer to GPU (upload the host mirror)
        """
        self._gpu.flush([self])

    def _ensure_host(self):
        if self._host is None or self._host.size != self.buffer.n:
            self._host = _host_empty(self.buffer.n, self._dtype)
            self._host_valid = False

    def _sync_host(self) -> np.ndarray:
        self.wait()
        self._ensure_host()
        if not self._host_valid:
            if self.buffer.nbytes:
                call("vkh_download", self._gpu._handle, self._host.ctypes.data,
                     self._ptr, self.buffer.nbytes)
            self._host_valid = True
        return self._host

    def _upload_host(self):
        if self._host is None or not self.buffer.nbytes:
            return
        call("vkh_upload", self._gpu._handle, self._ptr, self._host.ctypes.data, self.buffer.nbytes)
        self.job = self._gpu._job()
        self._upload_pending = True
        self._host_valid = True

    @property
    def array(self) -> np.ndarray:
        """Host view (synchronised on access).  Call ``flush()`` after writing it."""
        return self._sync_host().reshape(self.shape)

    def _info(self):
        return self.buffer

    def __getitem__(self, key: KeyType) -> ValueType:
        return self.array[key]

    def __setitem__(self, key: KeyType, value: ValueType):
        # The reference writes mapped memory without waiting (vkarray.py:176-177);
        # here the pending job is waited for first.
        if self._is_full_key(key) and np.ndim(value) == 0:
            self.wait()
            self._fill(value)
            return
        host = self._sync_host().reshape(self.shape)
        host[key] = value
        self._upload_host()

    def _is_full_key(self, key) -> bool:
        if key is Ellipsis:
            return True
        if isinstance(key, slice):
            return key == slice(None) and len(self.shape) >= 1
        return False

    def _fill(self, value):
        raise NotImplementedError

    def __repr__(self) -> str:
        return f"<{self.__class__.__name__}(shape={tuple(self.shape)})>"

    def __str__(self) -> str:
        return str(self.array)

    def __array__(self, dtype=None, copy=None) -> np.ndarray:
        a = self.array
        if dtype is not None:
            a = a.astype(dtype, copy=False)
        if copy:
            a = a.copy()
        return a

    def __len__(self):
        if len(self.shape) == 0:
            raise TypeError("len() of unsized object")
        return self.shape[0]

    @property
    def ndim(self) -> int:
        return len(self.shape)


class U32Array(_GPUArray):
    """
    GPU Array of uint (32bit) for shape or indices

    See Also
    --------
    vulkpy_amd.Array : float (32bit) array supporing mathematical operation

    Warnings
    --------
    U32Array doesn't support any mathematical operation.
    This class is designed for indexing and class label.
    """
    _dtype = np.uint32

    def __init__(self, gpu: GPU, *,
                 data: Optional[Iterable[int]] = None,
                 shape: Optional[Iterable[int]] = None):
        """
        Initialize U32Array

        Parameters
        ----------
        data : iterable of ints, optional
            Data to be copied.
        shape : iterable of ints, optional
            Shape for uninitialized array.
            If ``data`` is not ``None``, ``shape`` is ignored.

        Raises
        ------
        ValueError
            If neither ``data`` nor ``shape`` is specified.
        """
        super().__init__(gpu)

        if data is not None:
            self._init_data(data, np.uint32)
        else:
            if shape is None:
                raise ValueError("One of `data` or `shape` must be specified.")
            self._init_shape(shape, np.uint32)

    def _fill(self, value):
        call("vkh_fill_u32", self._gpu._handle, self._ptr, int(value) & 0xFFFFFFFF, self.buffer.n)
        self._set_job()

    def to_onehot(self, num_classes: int) -> "Array":
        """
        Convert to one hot vector

        Parameters
        ----------
        num_classes : int
            Number of classes

        Returns
        -------
        vulkpy_amd.Array
            One hot vector

        Notes
        -----
        The reference gathers rows of ``np.identity(num_classes)``
        (vkarray.py:239-253); here one kernel writes them, with labels beyond
        the last class clamped to it as that gather does.
        """
        num_classes = int(num_classes)
        if num_classes < 1:
            raise ValueError("num_classes must be positive")
        ret = Array(self._gpu, shape=tuple(self.shape) + (num_classes,))
        call("vkh_onehot", self._gpu._handle, self._ptr, self.size, num_classes, ret._ptr)
        ret._set_job([self])
        return ret


class Shape(U32Array):
    """
    GPU Array of uint (32bit) for shape
    """
    def __init__(self, gpu: GPU, *,
                 data: Optional[Iterable[int]] = None,
                 ndim: Optional[int] = None):
        """
        Initialize Shape

        Parameters
        ----------
        gpu : vulkpy_amd.GPU
            GPU
        data : iterable of ints, optional
            Data
        ndim : int, optional
            ndim
        """
        super().__init__(gpu, data=data, shape=(ndim,) if ndim is not None else None)


def _normalize_axes(axis, ndim: int) -> List[int]:
    """Flattened, unique, non-negative axes (NumPy rule; vkarray.py:1196)."""
    axes = np.unique(np.asarray(axis, dtype=int).reshape(-1))
    out = []
    for a in axes:
        a = int(a)
        if a < -ndim or a >= ndim:
            raise ValueError(f"axis {a} is out of bounds for array of dimension {ndim}")
        out.append(a + ndim if a < 0 else a)
    return sorted(set(out))


def _axis_runs(axes: List[int]) -> List[Tuple[int, int]]:
    """Adjacent axes merged into [first, last] runs, last run first."""
    runs: List[Tuple[int, int]] = []
    for a in axes:
        if runs and runs[-1][1] == a - 1:
            runs[-1] = (runs[-1][0], a)
        else:
            runs.append((a, a))
    return runs[::-1]


def _prod(s) -> int:
    return int(np.prod(s, dtype=np.int64)) if len(s) else 1


class Array(_GPUArray):
    """
    GPU Array for float (32bit)
    """

    def __init__(self, gpu: GPU, *,
                 data=None,
                 shape: Optional[Iterable[int]] = None):
        """
        Initialize Array

        Parameters
        ----------
        gpu : GPU
            GPU instance to allocate at.
        data : array_like, optional
            Data which copy to GPU buffer.
        shape : array_like, optional
            Array shape

        Raises
        ------
        ValueError
            If both ``data`` and ``shape`` are ``None``.
        """
        super().__init__(gpu)

        if data is not None:
            self._init_data(data, np.float32)
        elif shape is not None:
            self._init_shape(shape, np.float32)
        else:
            raise ValueError(f"`data` or `shape` must not be `None`.")

    def _fill(self, value):
        call("vkh_fill_f32", self._gpu._handle, self._ptr, float(value), self.buffer.n)
        self._set_job()

    def _new(self, shape) -> "Array":
        return Array(self._gpu, shape=shape)

    def _check_shape(self, other):
        if tuple(self.shape) != tuple(other.shape):
            raise ValueError(f"Incompatible shapes: {self.shape} vs {other.shape}")

    # ---- binary ops (F1-F6) -----------------------------------------------------
    def _op(self, other, op: str) -> "Array":
        h = self._gpu._handle
        if not isinstance(other, Array):
            s = float(other)
            ret = self._new(self.shape)
            call("vkh_binary_scalar", h, BINOP[op], self._ptr, s, ret._ptr, self.size)
            ret._set_job([self])
            return ret
        if tuple(self.shape) == tuple(other.shape):
            ret = self._new(self.shape)
            call("vkh_binary", h, BINOP[op], self._ptr, other._ptr, ret._ptr, self.size)
            ret._set_job([self, other])
            return ret
        shape = tuple(np.broadcast_shapes(tuple(self.shape), tuple(other.shape)))
        ret = self._new(shape)
        nd = len(shape)
        sa = (1,) * (nd - len(self.shape)) + tuple(self.shape)
        sb = (1,) * (nd - len(other.shape)) + tuple(other.shape)
        call("vkh_binary_bcast", h, BINOP[op], self._ptr, other._ptr, ret._ptr, nd,
             shape_array(sa), shape_array(sb), shape_array(shape))
        ret._set_job([self, other])
        return ret

    def _iop(self, other, op: str) -> "Array":
        h = self._gpu._handle
        if not isinstance(other, Array):
            call("vkh_binary_scalar", h, BINOP[op], self._ptr, float(other), self._ptr, self.size)
            self._set_job()
        elif tuple(self.shape) == tuple(other.shape):
            call("vkh_binary", h, BINOP[op], self._ptr, other._ptr, self._ptr, self.size)
            self._set_job([other])
        else:
            shape = tuple(np.broadcast_shapes(tuple(self.shape), tuple(other.shape)))
            if shape != tuple(self.shape):
                raise ValueError(f"Incompatible shape. {shape} vs {self.shape}")
            # NumPy rule (len(shape)) instead of the reference's shape[0]
            # (vkarray.py:542), which rejects most in-place broadcasts.
            nd = len(shape)
            sb = (1,) * (nd - len(other.shape)) + tuple(other.shape)
            call("vkh_binary_bcast", h, BINOP[op], self._ptr, other._ptr, self._ptr, nd,
                 shape_array(shape), shape_array(sb), shape_array(shape))
            self._set_job([other])
        return self

    def __add__(self, other: Union["Array", float]) -> "Array":
        return self._op(other, "add")

    def __sub__(self, other: Union["Array", float]) -> "Array":
        return self._op(other, "sub")

    def __mul__(self, other: Union["Array", float]) -> "Array":
        return self._op(other, "mul")

    def __truediv__(self, other: Union["Array", float]) -> "Array":
        return self._op(other, "div")

    def __iadd__(self, other: Union["Array", float]) -> "Array":
        return self._iop(other, "add")

    def __isub__(self, other: Union["Array", float]) -> "Array":
        return self._iop(other, "sub")

    def __imul__(self, other: Union["Array", float]) -> "Array":
        return self._iop(other, "mul")

    def __itruediv__(self, other: Union["Array", float]) -> "Array":
        return self._iop(other, "div")

    def _rop(self, other, op: str) -> "Array":
        ret = self._new(self.shape)
        call("vkh_binary_scalar", self._gpu._handle, BINOP[op], self._ptr, float(other),
             ret._ptr, self.size)
        ret._set_job([self])
        return ret

    def __radd__(self, other: float) -> "Array":
        return self._rop(other, "add")

    def __rsub__(self, other: float) -> "Array":
        return self._rop(other, "rsub")

    def __rmul__(self, other: float) -> "Array":
        return self._rop(other, "mul")

    def __rtruediv__(self, other: float) -> "Array":
        return self._rop(other, "rdiv")

    def __pow__(self, other: Union["Array", float]) -> "Array":
        return self._op(other, "pow")

    def __ipow__(self, other: Union["Array", float]) -> "Array":
        return self._iop(other, "pow")

    def __rpow__(self, other: float) -> "Array":
        return self._rop(other, "rpow")

    # ---- matmul (F15) ---------------------------------------------------------------
    def __matmul__(self, other: "Array") -> "Array":
        return self.matmul(other)

    def matmul(self, other: "Array", *, precision: Optional[str] = None) -> "Array":
        """
        Matrix product on the matrix cores

        Parameters
        ----------
        other : vulkpy_amd.Array
            Right operand; ``self.shape[-1]`` must equal ``other.shape[0]``.
        precision : {"fp32", "bf16", "bf16x3"}, optional
            Product precision (FP32 accumulation in every mode).  Default is
            ``vulkpy_amd.get_matmul_precision()`` (``"fp32"``).

        Notes
        -----
        Same shape rule as vkarray.py:585-605: the result has shape
        ``self.shape[:-1] + other.shape[1:]`` (``(1,)`` for 1-D @ 1-D).  For
        3-D operands the product contracts the last axis of ``self`` with the
        first axis of ``other`` (``numpy.tensordot(self, other, axes=1)``).
        """
        if ((len(self.shape) > 3) or (len(other.shape) > 3) or
                (self.shape[-1] != other.shape[0])):
            raise ValueError(f"Incompatible shapes: {self.shape} vs {other.shape}")
        shape = tuple(self.shape)[:-1] + tuple(other.shape)[1:]
        if len(shape) == 0:
            shape = (1,)
        M = _prod(self.shape[:-1])
        K = int(self.shape[-1])
        N = _prod(other.shape[1:])
        ret = self._new(shape)
        mode = GEMM_MODE[precision or _matmul_precision]
        call("vkh_gemm", self._gpu._handle, mode, 0, 0, M, N, K, 1.0,
             self._ptr, K, other._ptr, N, 0.0, ret._ptr, N, None, 0, None)
        ret._set_job([self, other])
        return ret

    def reshape(self, shape: Iterable[int]):
        """
        Reshape of this array (in place, like vkarray.py:607-622)

        Parameters
        ----------
        shape : iterable of int
            New shape

        Raises
        ------
        ValueError
            If ``shape`` is incompatible
        """
        self.shape = _reshape_target(self.buffer.n, shape)

    # ---- max / min -----------------------------------------------------------------------
    def max(self, other: Union["Array", float], inplace: bool = False) -> "Array":
        """
        Element-wise Max (numpy.maximum)

        Parameters
        ----------
        other : Array or float
            Other value
        inplace : bool
            If ``True``, update inplace, otherwise returns new array.
            Default value is ``False``.
        """
        if not inplace:
            return self._op(other, "max")
        return self._iop(other, "max")

    def min(self, other: Union["Array", float], inplace: bool = False) -> "Array":
        """
        Element-wise Min (numpy.minimum)

        Parameters
        ----------
        other : Array or float
            Other value
        inplace : bool
            If ``True``, update inplace, otherwise returns new array.
            Default value is ``False``.
        """
        if not inplace:
            return self._op(other, "min")
        return self._iop(other, "min")

    # ---- unary math (F7/F8) ----------------------------------------------------------------
    def _unary(self, fn: str, inplace: bool) -> "Array":
        if inplace:
            call("vkh_unary", self._gpu._handle, UNOP[fn], self._ptr, self._ptr, self.size)
            self._set_job()
            return self
        ret = self._new(self.shape)
        call("vkh_unary", self._gpu._handle, UNOP[fn], self._ptr, ret._ptr, self.size)
        ret._set_job([self])
        return ret

    def abs(self, inplace: bool = False) -> "Array":
        """Element-wise Abs"""
        return self._unary("abs", inplace)

    def sign(self, inplace: bool = False) -> "Array":
        """Element-wise sign"""
        return self._unary("sign", inplace)

    def sin(self, inplace: bool = False) -> "Array":
        """Element-wise sin()"""
        return self._unary("sin", inplace)

    def cos(self, inplace: bool = False) -> "Array":
        """Element-wise cos()"""
        return self._unary("cos", inplace)

    def tan(self, inplace: bool = False) -> "Array":
        """Element-wise tan()"""
        return self._unary("tan", inplace)

    def asin(self, inplace: bool = False) -> "Array":
        """Element-wise asin()"""
        return self._unary("asin", inplace)

    def acos(self, inplace: bool = False) -> "Array":
        """Element-wise acos()"""
        return self._unary("acos", inplace)

    def atan(self, inplace: bool = False) -> "Array":
        """Element-wise atan()"""
        return self._unary("atan", inplace)

    def sinh(self, inplace: bool = False) -> "Array":
        """Element-wise sinh()"""
        return self._unary("sinh", inplace)

    def cosh(self, inplace: bool = False) -> "Array":
        """Element-wise cosh()"""
        return self._unary("cosh", inplace)

    def tanh(self, inplace: bool = False) -> "Array":
        """Element-wise tanh()"""
        return self._unary("tanh", inplace)

    def asinh(self, inplace: bool = False) -> "Array":
        """Element-wise asinh()"""
        return self._unary("asinh", inplace)

    def acosh(self, inplace: bool = False) -> "Array":
        """Element-wise acosh()"""
        return self._unary("acosh", inplace)

    def atanh(self, inplace: bool = False) -> "Array":
        """Element-wise atanh()"""
        return self._unary("atanh", inplace)

    def exp(self, inplace: bool = False) -> "Array":
        """Element-wise exp()"""
        return self._unary("exp", inplace)

    def log(self, inplace: bool = False) -> "Array":
        """Element-wise log()"""
        return self._unary("log", inplace)

    def exp2(self, inplace: bool = False) -> "Array":
        """Element-wise exp2()"""
        return self._unary("exp2", inplace)

    def log2(self, inplace: bool = False) -> "Array":
        """Element-wise log2()"""
        return self._unary("log2", inplace)

    def sqrt(self, inplace: bool = False) -> "Array":
        """Element-wise sqrt()"""
        return self._unary("sqrt", inplace)

    def invsqrt(self, inplace: bool = False) -> "Array":
        """Element-wise 1/sqrt()"""
        return self._unary("invsqrt", inplace)

    # ---- clamp (F9) -------------------------------------------------------------------------
    def clamp(self, min: Union["Array", float], max: Union["Array", float],
              inplace: bool = False) -> "Array":
        """
        Element-wise clamp()

        Parameters
        ----------
        min : Array or float
            Minimum value
        max : Array or float
            Maximum value
        inplace : bool
            If ``True``, update inplace, otherwise returns new array.
            Default value is ``False``.

        Notes
        -----
        Array bounds are broadcast inside the kernel; the reference
        materialises them with ``broadcast_to`` first (vkarray.py:1147-1160).
        """
        lo_arr = isinstance(min, Array)
        hi_arr = isinstance(max, Array)
        shapes = [tuple(self.shape)]
        if lo_arr:
            shapes.append(tuple(min.shape))
        if hi_arr:
            shapes.append(tuple(max.shape))
        shape = tuple(np.broadcast_shapes(*shapes)) if len(shapes) > 1 else tuple(self.shape)
        if inplace and shape != tuple(self.shape):
            raise ValueError(f"Incompatible shape")
        lo_s = 0.0 if lo_arr else float(min)
        hi_s = 0.0 if hi_arr else float(max)
        nd = len(shape)

        def pad(s):
            return shape_array((1,) * (nd - len(s)) + tuple(s))

        out = self if inplace else self._new(shape)
        call("vkh_clamp", self._gpu._handle, self._ptr,
             min._ptr if lo_arr else None, max._ptr if hi_arr else None,
             lo_s, hi_s, out._ptr, nd, pad(self.shape),
             pad(min.shape) if lo_arr else None, pad(max.shape) if hi_arr else None,
             shape_array(shape))
        keep = [a for a, f in ((self, not inplace), (min, lo_arr), (max, hi_arr)) if f]
        out._set_job(keep)
        return out

    # ---- reductions (F11-F14) ------------------------------------------------------------------
    def _reduce(self, op: str, axis, keepdims: bool, rebroadcast: bool,
                mean: bool = False) -> "Array":
        h = self._gpu._handle
        ndim = len(self.shape)
        if rebroadcast:
            if not isinstance(axis, (int, np.integer)):
                raise ValueError("When `rebroadcast` is specified, " +
                                 "`axis` must be `int`")
            a = _normalize_axes(axis, ndim)[0]
            prev = _prod(self.shape[:a])
            K = int(self.shape[a])
            post = _prod(self.shape[a + 1:])
            ret = self._new(self.shape)
            call("vkh_reduce_axis", h, REDOP[op], self._ptr, ret._ptr, prev, K, post, 1,
                 1.0, float(K) if mean else 1.0)
            ret._set_job([self])
            return ret

        if axis is None:
            ret = self._new((1,))
            call("vkh_reduce_all", h, REDOP[op], self._ptr, self.size, ret._ptr, int(mean))
            ret._set_job([self])
            if keepdims:
                ret.reshape((1,) * ndim)
            return ret

        axes = _normalize_axes(axis, ndim)
        n_before = self.size
        runs = _axis_runs(axes)
        tmp: Array = self
        shape = list(self.shape)
        for i, (a0, a1) in enumerate(runs):
            prev = _prod(shape[:a0])
            K = _prod(shape[a0:a1 + 1])
            post = _prod(shape[a1 + 1:])
            new_shape = shape[:a0] + shape[a1 + 1:]
            ret = self._new(tuple(new_shape))
            mul = 1.0
            if mean and i == len(runs) - 1:
                # vkarray.py:1429-1430: ret *= n_after / n_before (FP32 scalar)
                mul = float(np.float32(_prod(new_shape) / n_before))
            call("vkh_reduce_axis", h, REDOP[op], tmp._ptr, ret._ptr, prev, K, post, 0, mul, 1.0)
            ret._set_job([tmp])
            tmp = ret
            shape = new_shape
        if keepdims:
            ks = list(self.shape)
            for a in axes:
                ks[a] = 1
            tmp.reshape(tuple(ks))
        return tmp

    def sum(self, axis: Union[None, int, Iterable[int]] = None,
            keepdims: bool = False,
            rebroadcast: bool = False) -> "Array":
        """
        Calculate Sum of Elements

        Parameters
        ----------
        axis : int, optional
            Reduction axis
        keepdims : bool, optional
            When `True`, reduced dimensions are keeped with size one.
            Default is `False`.
        rebroadcast : bool, optional
            When `True`, keep shape by re-broadcasting after reduce.
            Default is `False`.
        """
        return self._reduce("sum", axis, keepdims, rebroadcast)

    def prod(self, axis: Union[None, int, Iterable[int]] = None,
             keepdims: bool = False,
             rebroadcast: bool = False) -> "Array":
        """
        Calculate Product of Elements
        """
        return self._reduce("prod", axis, keepdims, rebroadcast)

    def maximum(self, axis: Union[None, int, Iterable[int]] = None,
                keepdims: bool = False,
                rebroadcast: bool = False) -> "Array":
        """
        Get Maximum Value
        """
        return self._reduce("maximum", axis, keepdims, rebroadcast)

    def minimum(self, axis: Union[None, int, Iterable[int]] = None,
                keepdims: bool = False,
                rebroadcast: bool = False) -> "Array":
        """
        Get Minimum Value
        """
        return self._reduce("minimum", axis, keepdims, rebroadcast)

    def mean(self, axis: Union[None, int, Iterable[int]] = None,
             keepdims: bool = False,
             rebroadcast: bool = False) -> "Array":
        """
        Calculate Mean Value

        Notes
        -----
        ``axis=None`` divides the FP64 sum by the element count before the one
        rounding to FP32 (the reference multiplies the FP32 sum by an FP32
        ``1/n``, vkarray.py:1429-1430).  Axis means keep the reference's FP32
        scaling.
        """
        return self._reduce("sum", axis, keepdims, rebroadcast, mean=True)

    # ---- broadcast / gather (F10, F17) ----------------------------------------------------------
    def broadcast_to(self, shape: Iterable[int]) -> "Array":
        """
        Broadcast to new shape

        Raises
        ------
        ValueError
            If ``shape`` is not compatible (NumPy rule).
        """
        shape = _parse_shape(shape)
        if tuple(np.broadcast_shapes(tuple(self.shape), shape)) != shape:
            raise ValueError(f"Cannot broadcast to {shape}")
        ret = self._new(shape)
        nd = len(shape)
        sa = (1,) * (nd - len(self.shape)) + tuple(self.shape)
        call("vkh_broadcast_to", self._gpu._handle, self._ptr, ret._ptr, nd,
             shape_array(sa), shape_array(shape))
        ret._set_job([self])
        return ret

    def gather(self, indices: U32Array, axis: Optional[int] = None) -> "Array":
        """
        Gather values of indices

        Parameters
        ----------
        indices : vulkpy_amd.U32Array
            Indices
        axis : int, optional
            Axis of gather.
            If ``None`` (default), array is flattened beforehand.

        Notes
        -----
        Out-of-range indices read the last element (flat) or the last slice
        along ``axis``.  The output of an axis gather has shape
        ``indices.shape + shape[:axis] + shape[axis+1:]`` (vkarray.py:1509).
        """
        h = self._gpu._handle
        n = indices.size
        if axis is None:
            ret = self._new(indices.shape)
            call("vkh_gather", h, self._ptr, self.size, indices._ptr, ret._ptr, n)
        else:
            a = _normalize_axes(axis, len(self.shape))[0]
            prev_shape = tuple(self.shape[:a])
            post_shape = tuple(self.shape[a + 1:])
            ret = self._new(tuple(indices.shape) + prev_shape + post_shape)
            call("vkh_gather_axis", h, self._ptr, indices._ptr, ret._ptr,
                 _prod(prev_shape), int(self.shape[a]), _prod(post_shape), n)
        ret._set_job([self, indices])
        return ret


def zeros(gpu: GPU, shape: Iterable[int]) -> Array:
    """
    Zero initialized Array

    Parameters
    ----------
    gpu : vulkpy_amd.GPU
        GPU
    shape : iterable of ints
        Shape

    Returns
    -------
    vulkpy_amd.Array
        Zero initialized Array
    """
    z = Array(gpu, shape=shape)
    z._fill(0.0)
    return z


_matmul_precision = "fp32"


def set_matmul_precision(precision: str):
    """Default precision of ``@`` / ``Array.matmul``: "fp32", "bf16" or "bf16x3"."""
    global _matmul_precision
    if precision not in GEMM_MODE:
        raise ValueError(f"unknown precision {precision!r}")
    _matmul_precision = precision


def get_matmul_precision() -> str:
    return _matmul_precision


def gemm_configs(precision: str) -> List[dict]:
    """Kernel configurations of the matrix products in a precision mode:
    tile ``bm`` x ``bn``, k slice ``bk``, waves per block and ``direct``
    (operand slices loaded from global memory straight into LDS)."""
    mode = GEMM_MODE[precision]
    n = ctypes.c_int()
    call("vkh_gemm_configs", mode, ctypes.byref(n))
    out = []
    for i in range(n.value):
        v = [ctypes.c_int() for _ in range(5)]
        call("vkh_gemm_config_info", mode, i, *[ctypes.byref(x) for x in v])
        out.append({"bm": v[0].value, "bn": v[1].value, "bk": v[2].value, "waves": v[3].value,
                    "direct": bool(v[4].value)})
    return out


def autotune_gemm(gpu: GPU, M: int, N: int, K: int, *, precision: str = "bf16",
                  trans_a: bool = False, trans_b: bool = False, reps: int = 3) -> dict:
    """
    Time every kernel configuration of an ``op(A)[M,K] @ op(B)[K,N]`` product
    on ``gpu`` and keep the fastest for that shape (later products of the
    same shape, layout and precision use it).

    Returns
    -------
    dict
        ``best`` (index into :func:`gemm_configs`) and ``ms`` (mean time of
        each configuration, operand packing included; -1 where the shape
        has a single kernel choice).
    """
    mode = GEMM_MODE[precision]
    n = len(gemm_configs(precision))
    ms = (ctypes.c_float * n)()
    best = ctypes.c_int()
    call("vkh_gemm_autotune", gpu._handle, mode, int(trans_a), int(trans_b), int(M), int(N), int(K),
         int(reps), ctypes.byref(best), ms)
    return {"best": best.value, "ms": [round(float(x), 5) for x in ms]}


REDUCE_FAMILY = {-1: None, 0: "all", 1: "rows", 2: "cols"}


def reduce_configs(family: str) -> List[dict]:
    """Kernel configurations of a reduction family (``"all"``, ``"rows"`` or
    ``"cols"``): ``unroll`` = 128-bit loads in flight per thread and ``par`` =
    blocks per CU (``all``, ``rows``) or target threads per CU (``cols``,
    which sets how many chunks the reduced axis is split into)."""
    fam = {v: k for k, v in REDUCE_FAMILY.items()}[family]
    n = ctypes.c_int()
    call("vkh_reduce_configs", fam, ctypes.byref(n))
    out = []
    for i in range(n.value):
        u, p = ctypes.c_int(), ctypes.c_int()
        call("vkh_reduce_config_info", fam, i, ctypes.byref(u), ctypes.byref(p))
        out.append({"unroll": u.value, "par": p.value})
    return out


def _reduce_problems(shape, axis) -> List[Tuple[int, int, int, int]]:
    """(kind, prev, K, post) of the kernel calls ``Array._reduce`` makes for a
    reduction of an array of ``shape`` over ``axis`` (without rebroadcast)."""
    shape = tuple(int(s) for s in shape)
    if axis is None:
        return [(0, 1, _prod(shape), 1)]
    out = []
    dims = list(shape)
    for a0, a1 in _axis_runs(_normalize_axes(axis, len(shape))):
        out.append((1, _prod(dims[:a0]), _prod(dims[a0:a1 + 1]), _prod(dims[a1 + 1:])))
        dims = dims[:a0] + dims[a1 + 1:]
    return out


def autotune_reduce(gpu: GPU, shape, axis=None, *, reps: int = 3,
                    same_order: bool = False) -> List[dict]:
    """
    Time every kernel configuration of the reduction of an array of ``shape``
    over ``axis`` (``None``, an int or a tuple of ints, as in
    :meth:`Array.sum`) on ``gpu`` and keep the fastest for each kernel call
    it makes; later reductions of that shape (any op, eager or captured) use
    it.  With ``same_order`` only the configurations that keep the default
    summation order are timed (``ms`` is -1 for the others), so the choice
    cannot change any result.

    Returns
    -------
    list of dict
        one per kernel call: ``family`` (None where there is a single kernel
        choice), ``best`` (index into :func:`reduce_configs`), ``config`` and
        ``ms`` (mean time of each configuration).
    """
    out = []
    for kind, prev, K, post in _reduce_problems(shape, axis):
        if prev * K * post == 0:
            continue
        f = ctypes.c_int()
        call("vkh_reduce_family", kind, prev, K, post, ctypes.byref(f))
        n = ctypes.c_int()
        call("vkh_reduce_configs", f.value, ctypes.byref(n))
        ms = (ctypes.c_float * n.value)()
        best = ctypes.c_int()
        call("vkh_reduce_autotune", gpu._handle, kind, prev, K, post, int(reps), int(same_order),
             ctypes.byref(best), ms)
        fam = REDUCE_FAMILY[f.value]
        out.append({"family": fam, "shape": (kind, prev, K, post), "best": best.value,
                    "config": reduce_configs(fam)[best.value] if fam else None,
                    "ms": [round(float(x), 5) for x in ms]})
    return out
