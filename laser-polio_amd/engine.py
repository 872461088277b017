"""ctypes binding of liblpstep.so (include/lp_step.h) and its device buffers.

The engine owns one torch tensor per bound buffer (torch is used only as the
device allocator and for host<->device copies) and one library context.  It
raises loudly when the HIP library or a GPU is missing: there is no CPU path.
"""

import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liblpstep.so")

LP_I8, LP_U8, LP_I16, LP_U16, LP_I32, LP_U32, LP_I64, LP_U64, LP_F32, LP_F64 = range(1, 11)
_DTYPE_CODE = {
    np.dtype(np.int8): LP_I8,
    np.dtype(np.uint8): LP_U8,
    np.dtype(np.int16): LP_I16,
    np.dtype(np.uint16): LP_U16,
    np.dtype(np.int32): LP_I32,
    np.dtype(np.uint32): LP_U32,
    np.dtype(np.int64): LP_I64,
    np.dtype(np.uint64): LP_U64,
    np.dtype(np.float32): LP_F32,
    np.dtype(np.float64): LP_F64,
}
# torch has no general uint16/uint32/uint64 kernels: keep those buffers as signed views.
_TORCH_VIEW = {np.dtype(np.uint16): np.int16, np.dtype(np.uint32): np.int32, np.dtype(np.uint64): np.int64}

STAGE_VITAL, STAGE_DISEASE, STAGE_RI, STAGE_SIA, STAGE_TRANSMISSION, STAGE_LOG = (1, 2, 4, 8, 16, 32)
STAGE_ALL = 0x3F
DAY_VD_STEP, DAY_VD_COPY, DAY_DS, DAY_RI, DAY_SIA, DAY_TX, DAY_STOPCHK = (1, 2, 4, 8, 16, 32, 64)
TILE = 1024
CHUNK = 256
MAX_STRAINS = 4
KM_LEN = 101

ERRORS = {-1: "bad argument", -2: "HIP error", -3: "node capacity exhausted", -4: "call order", -5: "unbound buffer", -6: "overflow"}


class EngineUnavailable(RuntimeError):
    pass


class LPError(RuntimeError):
    pass


class LPConfig(ctypes.Structure):
    _fields_ = [
        ("capacity", ctypes.c_int64),
        ("n_nodes", ctypes.c_int32),
        ("n_sims", ctypes.c_int32),
        ("n_strains", ctypes.c_int32),
        ("nt", ctypes.c_int32),
        ("n_sia_events", ctypes.c_int32),
        ("tiles_per_block", ctypes.c_int32),
        ("mix_mode", ctypes.c_int32),
        ("ei_grid", ctypes.c_int32),
        ("sample_grid", ctypes.c_int32),
        ("commit_grid", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 4),
    ]


SIM_PARAMS_DTYPE = np.dtype(
    [
        ("seed", np.uint64),
        ("slot_base", np.int64),
        ("net_off", np.int64),
        ("node_begin", np.int32),
        ("n_nodes", np.int32),
        ("strain_r0", np.float32, (MAX_STRAINS,)),
        ("thr_paralysis", np.uint64),
        ("thr_zi", np.uint64),
        ("zi", np.float32),
        ("disp_r", np.int32),
        ("ri_strain", np.int32),
        ("ri_step", np.int32),
        ("newborn_ri_timer", np.int32),
        ("last_seed_day", np.int32),
    ],
    align=True,
)

LP_DIST_CONSTANT, LP_DIST_NORMAL, LP_DIST_GAMMA, LP_DIST_LOGNORMAL, LP_DIST_POISSON, LP_DIST_EXPONENTIAL, \
    LP_DIST_UNIFORM_INT = range(7)
INIT_BANDS = 8
INIT_PYR = 32


class LPDist(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("a", ctypes.c_float), ("b", ctypes.c_float)]


class LPGravitySpec(ctypes.Structure):
    """lp_gravity_spec of include/lp_step.h."""
    _fields_ = [("k", ctypes.c_double), ("a", ctypes.c_double), ("b", ctypes.c_double), ("c", ctypes.c_double),
                ("max_rowsum", ctypes.c_double), ("reserved", ctypes.c_int32 * 2)]


class LPInitSim(ctypes.Structure):
    """lp_init_sim of include/lp_step.h: the device initialisation of one simulation."""

    _fields_ = [
        ("seed", ctypes.c_uint64),
        ("age_model", ctypes.c_int32),
        ("age_mean_days", ctypes.c_float),
        ("age_cap_days", ctypes.c_float),
        ("n_pyr", ctypes.c_int32),
        ("pyr_cdf", ctypes.c_float * INIT_PYR),
        ("pyr_lo", ctypes.c_int32 * INIT_PYR),
        ("pyr_hi", ctypes.c_int32 * INIT_PYR),
        ("init_immun", ctypes.c_float),
        ("n_bands", ctypes.c_int32),
        ("band_lo", ctypes.c_int32 * INIT_BANDS),
        ("band_hi", ctypes.c_int32 * INIT_BANDS),
        ("band_frac", ctypes.c_float * INIT_BANDS),
        ("dur_exp", LPDist),
        ("dur_inf", LPDist),
        ("t_to_paralysis", LPDist),
        ("het_model", ctypes.c_int32),
        ("risk_mu", ctypes.c_float),
        ("risk_sigma", ctypes.c_float),
        ("inf_mean", ctypes.c_float),
        ("rho", ctypes.c_float),
        ("ri", ctypes.c_int32),
        ("ri_lo", ctypes.c_float),
        ("ri_hi", ctypes.c_float),
        ("vd", ctypes.c_int32),
        ("missed_frac", ctypes.c_float),
        ("reserved", ctypes.c_int32 * 3),
    ]


_lib = None


def load_library(path=LIB_PATH):
    """Load liblpstep.so once and declare its C signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise EngineUnavailable(f"{path} is not built; run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    lib = ctypes.CDLL(path)
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    lib.lp_abi_version.restype = i32
    lib.lp_sizeof_sim_params.restype = i64
    lib.lp_create.argtypes = [ctypes.POINTER(LPConfig), ctypes.POINTER(vp)]
    lib.lp_bind.argtypes = [vp, ctypes.c_char_p, vp, i32, i64]
    lib.lp_bind_names.restype = ctypes.c_char_p
    lib.lp_prepare.argtypes = [vp]
    lib.lp_set_stream.argtypes = [vp, vp]
    lib.lp_owns_stream.argtypes = [vp]
    lib.lp_run_days.argtypes = [vp, i32, i32, i32]
    lib.lp_step_stage.argtypes = [vp, i32, i32]
    for f in ("lp_step_vital", "lp_step_disease", "lp_step_ri", "lp_step_sia", "lp_step_transmission", "lp_log"):
        getattr(lib, f).argtypes = [vp, i32]
    lib.lp_sync.argtypes = [vp]
    lib.lp_invalidate.argtypes = [vp]
    lib.lp_stop_days.argtypes = [vp, ctypes.POINTER(i32), i32]
    lib.lp_debug_buffer.argtypes = [vp, ctypes.c_char_p]
    lib.lp_debug_buffer.restype = vp
    lib.lp_set_timing.argtypes = [vp, i32]
    lib.lp_agent_pass_timing.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(ctypes.c_double)]
    lib.lp_agent_pass_launch_ms.argtypes = [vp, ctypes.POINTER(ctypes.c_float), i64]
    lib.lp_agent_pass_launch_ms.restype = i64
    lib.lp_sizeof_init_sim.restype = i64
    lib.lp_init_population.argtypes = [vp, ctypes.POINTER(LPInitSim), i32, vp, i64]
    lib.lp_build_network.argtypes = [vp, ctypes.POINTER(LPGravitySpec), i32, vp, vp]
    lib.lp_last_error.argtypes = [vp]
    lib.lp_last_error.restype = ctypes.c_char_p
    lib.lp_destroy.argtypes = [vp]
    lib.lp_destroy.restype = None
    if lib.lp_abi_version() != 1:
        raise EngineUnavailable("liblpstep.so ABI version mismatch")
    if lib.lp_sizeof_sim_params() != SIM_PARAMS_DTYPE.itemsize:
        raise EngineUnavailable("lp_sim_params layout mismatch between include/lp_step.h and engine.py")
    if lib.lp_sizeof_init_sim() != ctypes.sizeof(LPInitSim):
        raise EngineUnavailable("lp_init_sim layout mismatch between include/lp_step.h and engine.py")
    _lib = lib
    return lib


def _require_gpu():
    import torch

    if not torch.cuda.is_available():
        raise EngineUnavailable("no GPU visible: the laser-polio step runs only on the HIP path (MI355X, gfx950)")
    return torch


def thr_from_prob64(p):
    """ceil(p * 2^32) clamped to [0, 2^32], from float64 probabilities (host-side thresholds)."""
    p = np.asarray(p, dtype=np.float64)
    out = np.ceil(np.clip(p, 0.0, 1.0) * 4294967296.0)
    out = np.where(p > 0, out, 0.0)
    return out.astype(np.uint64)


class Engine:
    """One library context plus the device copies of every bound buffer."""

    def __init__(self, config, arrays, timing=False, device_only=()):
        """device_only: names of buffers allocated on the device without uploading the host
        array (columns that lp_init_population writes)."""
        self.torch = _require_gpu()
        self.lib = load_library()
        self.device = self.torch.device("cuda", self.torch.cuda.current_device())
        self.arrays = arrays
        self.tensors = {}
        self.cfg = config
        ctx = ctypes.c_void_p()
        rc = self.lib.lp_create(ctypes.byref(config), ctypes.byref(ctx))
        if rc != 0:
            raise LPError(f"lp_create failed ({ERRORS.get(rc, rc)}): check sizes {config.capacity=} {config.n_nodes=}")
        self.ctx = ctx
        for name, arr in arrays.items():
            if name in device_only:
                dt = np.dtype(_TORCH_VIEW.get(arr.dtype, arr.dtype))
                self.tensors[name] = self.torch.empty(arr.size, dtype=_torch_dtype(self.torch, dt), device=self.device)
            else:
                self._upload_new(name, arr)
        self.torch.cuda.synchronize()
        for name, arr in arrays.items():
            t = self.tensors[name]
            n = arr.size if name != "sim_params" else arr.nbytes
            code = LP_U8 if name == "sim_params" else _DTYPE_CODE[arr.dtype]
            self._check(self.lib.lp_bind(self.ctx, name.encode(), ctypes.c_void_p(t.data_ptr()), code, int(n)))
        self._check(self.lib.lp_prepare(self.ctx))
        if timing:
            self._check(self.lib.lp_set_timing(self.ctx, 1))

    # -- buffers -----------------------------------------------------------------
    def _upload_new(self, name, arr):
        torch = self.torch
        a = np.ascontiguousarray(arr)
        if name == "sim_params":
            a = a.view(np.uint8)
        elif a.dtype in _TORCH_VIEW:
            a = a.view(_TORCH_VIEW[a.dtype])
        self.tensors[name] = torch.from_numpy(a.reshape(-1)).to(self.device)

    def init_population(self, specs, infected_slots):
        """lp_init_population: the agent columns drawn on the device from `specs` (one
        LPInitSim per simulation); `infected_slots` (int64) start infectious."""
        arr = (LPInitSim * len(specs))(*specs)
        inf = np.ascontiguousarray(infected_slots, dtype=np.int64)
        self._check(self.lib.lp_init_population(self.ctx, arr, len(specs), ctypes.c_void_p(inf.ctypes.data if inf.size else 0),
                                                int(inf.size)))

    def push(self, name, arr):
        """Host -> device copy of one bound buffer (same size as at bind time)."""
        a = np.ascontiguousarray(arr)
        if name == "sim_params":
            a = a.view(np.uint8)
        elif a.dtype in _TORCH_VIEW:
            a = a.view(_TORCH_VIEW[a.dtype])
        self.sync()
        self.tensors[name].copy_(self.torch.from_numpy(a.reshape(-1)))
        self.torch.cuda.synchronize()
        self._check(self.lib.lp_invalidate(self.ctx))

    def pull(self, name, out):
        """Device -> host copy into `out` (numpy, same size and dtype as bound)."""
        self.sync()
        src = self.tensors[name].cpu().numpy()
        out.reshape(-1).view(src.dtype)[...] = src

    # -- stepping ------------------------------------------------------------------
    def _check(self, rc):
        if rc != 0:
            msg = self.lib.lp_last_error(self.ctx).decode() if self.ctx else ""
            raise LPError(f"liblpstep error {rc} ({ERRORS.get(rc, '?')}): {msg}")

    def set_stream(self, stream):
        """Queue the library's work on `stream` (a torch.cuda.Stream or a raw hipStream_t
        handle) instead of its own stream; the caller keeps ownership of it."""
        handle = getattr(stream, "cuda_stream", stream)
        self._check(self.lib.lp_set_stream(self.ctx, ctypes.c_void_p(int(handle) if handle else None)))

    def owns_stream(self):
        return bool(self.lib.lp_owns_stream(self.ctx))

    def run_days(self, t_begin, t_end, use_graph=True):
        self._check(self.lib.lp_run_days(self.ctx, int(t_begin), int(t_end), 1 if use_graph else 0))

    def step(self, t, mask):
        self._check(self.lib.lp_step_stage(self.ctx, int(t), int(mask)))

    def sync(self):
        self._check(self.lib.lp_sync(self.ctx))

    def stop_days(self):
        out = (ctypes.c_int32 * self.cfg.n_sims)()
        self._check(self.lib.lp_stop_days(self.ctx, out, self.cfg.n_sims))
        return np.array(list(out), dtype=np.int32)

    def set_timing(self, on):
        """On: lp_run_days launches each day's kernels directly and brackets its agent pass
        with HIP events (read by agent_pass_timing); off: days replay their hipGraphs."""
        self._check(self.lib.lp_set_timing(self.ctx, 1 if on else 0))

    def build_network(self, specs, pop, dist):
        """Every simulation's gravity network on the device (lp_build_network): `specs` =
        one (k, a, b, c, max_rowsum) per simulation, `pop` / `dist` = float64 device tensors
        (one population per node; each simulation's n x n distances at its network offset)."""
        import torch

        if pop.dtype != torch.float64 or dist.dtype != torch.float64 or not pop.is_cuda or not dist.is_cuda:
            raise ValueError("build_network: pop and dist must be float64 device tensors")
        if pop.numel() != self.cfg.n_nodes or dist.numel() != self.tensors["network_t"].numel():
            raise ValueError("build_network: pop needs one value per node, dist the size of network_t")
        arr = (LPGravitySpec * len(specs))()
        for i, (k, a, b, c, mx) in enumerate(specs):
            arr[i].k, arr[i].a, arr[i].b, arr[i].c, arr[i].max_rowsum = k, a, b, c, mx
        self._check(self.lib.lp_build_network(self.ctx, arr, len(specs), ctypes.c_void_p(pop.data_ptr()),
                                              ctypes.c_void_p(dist.data_ptr())))

    def agent_pass_launch_ms(self):
        """Milliseconds of each event-bracketed agent-pass launch since the last
        agent_pass_timing() call (does not reset)."""
        n = int(self.lib.lp_agent_pass_launch_ms(self.ctx, None, 0))
        if n < 0:
            raise LPError(f"lp_agent_pass_launch_ms: error {n}")
        out = np.zeros(n, np.float32)
        if n:
            got = int(self.lib.lp_agent_pass_launch_ms(self.ctx, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n))
            if got < 0:
                raise LPError(f"lp_agent_pass_launch_ms: error {got}")
        return out

    def agent_pass_timing(self):
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        self._check(self.lib.lp_agent_pass_timing(self.ctx, ctypes.byref(n), ctypes.byref(ms)))
        return int(n.value), float(ms.value)

    def debug(self, name, count, dtype):
        """Copy a library diagnostics buffer (see lp_debug_buffer) to the host."""
        self.sync()
        ptr = self.lib.lp_debug_buffer(self.ctx, name.encode())
        if not ptr:
            raise LPError(f"unknown debug buffer {name}")
        out = np.empty(count, dtype=dtype)
        self._memcpy_d2h(out, ptr)
        return out

    def _memcpy_d2h(self, out, dev_ptr):
        hiprt = _hip_runtime()
        rc = hiprt.hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(dev_ptr), ctypes.c_size_t(out.nbytes), 2)
        if rc != 0:
            raise LPError(f"hipMemcpy failed ({rc})")

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.lp_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 -- interpreter shutdown
            pass


def _torch_dtype(torch, dt):
    return {np.dtype(np.int8): torch.int8, np.dtype(np.uint8): torch.uint8, np.dtype(np.int16): torch.int16,
            np.dtype(np.int32): torch.int32, np.dtype(np.int64): torch.int64, np.dtype(np.float32): torch.float32,
            np.dtype(np.float64): torch.float64}[np.dtype(dt)]


_hiprt = None


def _hip_runtime():
    global _hiprt
    if _hiprt is None:
        for name in ("libamdhip64.so", "libamdhip64.so.7", "libamdhip64.so.6"):
            try:
                _hiprt = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _hiprt is None:
            raise EngineUnavailable("libamdhip64 not found")
        _hiprt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return _hiprt
