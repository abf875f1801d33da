"""Host-side mirror of the reference's walker-step interface over the C-ABI.

``AfqmcSys`` carries the method names of ``base::afqmc_sys``
(src/AFQMC/afqmc_sys.hpp:44-230) and the driver's free functions
(``get_vbias``, ``get_vHS``), with the reference's array layouts on the host:
walkers ``[nwalk][2][M][N]``, ``W_data [nwalk][8]``, compact G
``[2NM][nwalk]``, ``vbias``/``X`` ``[K][nwalk]``, ``vHS [M*M][nwalk]``.  The
arithmetic runs on the GPU through libafqmc_hip.so; nothing here computes.
"""
import ctypes as C

import numpy as np

from . import _native as nat
from .system import System

_D = C.POINTER(C.c_double)


def _dp(a):
    return a.view(np.float64).ctypes.data_as(_D) if a.dtype == np.complex128 else a.ctypes.data_as(_D)


# The reference driver's timers (src/miniapps/miniafqmc_base.cpp:50-77, the
# phases its step loop brackets at :248-350) and the library call sites whose
# GPU time each one covers.  "Total" is the wall time of the run; launches at
# upload and the tile autotuner's trial launches fall under no timer.
REFERENCE_TIMERS = (
    ("compact Mixed Density Matrix", ("gc_gemm", "det_restore")),
    ("Mixed Density Matrix", ("gfull_gemm",)),
    ("Bias Potential", ("vbias_gemm", "vbias_reduce", "csrmm_vbias", "csrmm_vbias_t")),
    ("H-S Potential", ("vhs_gemm", "csrmm_vhs", "vhs_transpose")),
    ("Sigma", ("fields", "set_substep")),
    ("Propagation", ("prop_pw", "prop_taylor", "prop_ps", "prop_split_w", "prop_fused")),
    ("Other", ("save_overlaps", "weights", "branch_select", "branch_gather")),
    ("Overlap", ("ovlp_gemm", "gj_inverse")),
    ("Orthgonalization", ("ortho_gram", "ortho_chol_inv", "ortho_sign", "ortho_apply", "ortho_flags",
                          "lq_orthogonalize", "lq_fallback")),
    ("Local Energy", ("energy_one", "energy_two", "energy_direct", "energy_exch_dot", "energy_pairsum",
                      "energy_assemble", "energy_zero", "energy_trace", "energy_v_gemm", "energy_v_reduce", "energy_vx_gemm",
                      "energy_y_gemm", "csrmm_vakbl", "g_transpose", "gc64_gemm", "convert", "g_split3")),
)


def reference_timers(stats):
    """Per-call-site kernel statistics ({name: (launches, ms)}, as
    AfqmcSys.kernel_stats returns) summed under the reference driver's timer
    names: {timer: (launches, ms)}.  Call sites under no timer are left out."""
    out = {}
    for timer, sites in REFERENCE_TIMERS:
        n = sum(stats[k][0] for k in sites if k in stats)
        ms = sum(stats[k][1] for k in sites if k in stats)
        out[timer] = (n, ms)
    return out


class AfqmcSys:
    def __init__(self, system: System, nwalk: int, mode: str = "fp64_csr", cutoff: float = 1e-6,
                 use_vakbl: bool = True, dense_exchange: bool = True, energy_precision: int = 1,
                 transposed: bool = True, three_piece: int = 0, csr_dense: bool = True,
                 coulomb_three_piece: bool = False):
        self.sys = system
        self.M, self.N, self.K, self.nwalk = system.nmo, system.naea, system.nchol, nwalk
        self.mode = mode
        h = C.c_void_p()
        nat.check(nat.lib.afqmc_ctx_create(self.M, self.N, self.K, nwalk, nat.MODES[mode], C.byref(h)))
        self.ctx = h
        self._check(nat.lib.afqmc_set_option(self.ctx, nat.OPT_DENSE_EXCHANGE, 1.0 if dense_exchange else 0.0))
        self._check(nat.lib.afqmc_set_option(self.ctx, nat.OPT_ENERGY_PRECISION, float(energy_precision)))
        self._check(nat.lib.afqmc_set_option(self.ctx, nat.OPT_TRANSPOSED, 1.0 if transposed else 0.0))
        if mode == "fp64_csr":
            # a half-full (or fuller) SpvnT also kept dense for the bias product
            self._check(nat.lib.afqmc_set_option(self.ctx, nat.OPT_CSR_DENSE, 1.0 if csr_dense else 0.0))
        if three_piece:
            # split-BF16 mode: contractions in three-piece split-BF16 (1 bias, 2 vHS, 4 propagator)
            self._check(nat.lib.afqmc_set_option(self.ctx, nat.OPT_THREE_PIECE, float(three_piece)))
        if coulomb_three_piece:
            # split-BF16 mode, energy level 2: the Coulomb term's contraction in three pieces
            self._check(nat.lib.afqmc_set_option(self.ctx, nat.OPT_COULOMB_THREE_PIECE, 1.0))
        ta = np.ascontiguousarray(system.trial_alpha)
        tb = np.ascontiguousarray(system.trial_beta)
        sp = system.spvn.view()
        vk = system.vakbl.view() if (use_vakbl and system.vakbl is not None) else None
        self._check(nat.lib.afqmc_upload_system(
            self.ctx, _dp(ta), _dp(tb), _dp(np.ascontiguousarray(system.propg1)),
            _dp(np.ascontiguousarray(system.haj)), C.byref(sp), C.byref(vk) if vk is not None else None,
            cutoff, system.dt))

    # -- plumbing -------------------------------------------------------------
    def _check(self, code):
        nat.check(code, self.ctx)

    def close(self):
        if self.ctx:
            nat.lib.afqmc_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        self._check(nat.lib.afqmc_sync(self.ctx))

    _shapes = {
        nat.BUF_W: lambda s: (s.nwalk, 2, s.M, s.N),
        nat.BUF_WDATA: lambda s: (s.nwalk, 8),
        nat.BUF_G: lambda s: (2 * s.N * s.M, s.nwalk),
        nat.BUF_VBIAS: lambda s: (s.K, s.nwalk),
        nat.BUF_X: lambda s: (s.K, s.nwalk),
        nat.BUF_VHS: lambda s: (s.M * s.M, s.nwalk),
        nat.BUF_HYBRIDW: lambda s: (s.nwalk,),
        nat.BUF_EPARTS: lambda s: (s.nwalk, 3),
        nat.BUF_GFULL: lambda s: (2 * s.M * s.M, s.nwalk),
    }

    def get(self, buf):
        out = np.empty(self._shapes[buf](self), np.complex128)
        self._check(nat.lib.afqmc_copy(self.ctx, buf, 0, _dp(out)))
        return out

    def set(self, buf, arr):
        a = np.ascontiguousarray(arr, dtype=np.complex128)
        assert a.shape == self._shapes[buf](self), (a.shape, self._shapes[buf](self))
        self._check(nat.lib.afqmc_copy(self.ctx, buf, 1, _dp(a)))

    @property
    def W(self):
        return self.get(nat.BUF_W)

    @property
    def W_data(self):
        return self.get(nat.BUF_WDATA)

    # -- reference interface ---------------------------------------------------
    def init_walkers(self):
        """Walkers = conj(trialwfn), weights 1 (miniafqmc_base.cpp:225-236)."""
        self._check(nat.lib.afqmc_init_walkers(self.ctx))

    def calculate_mixed_density_matrix(self, compact=True):
        """afqmc_sys::calculate_mixed_density_matrix (afqmc_sys.hpp:95-116); returns G [2NM][nw]."""
        if not compact:
            raise nat.AfqmcError(8, "full (non-compact) density matrix")
        self._check(nat.lib.afqmc_mixed_density_matrix(self.ctx))
        return self.get(nat.BUF_G)

    def calculate_overlaps(self):
        """afqmc_sys::calculate_overlaps (afqmc_sys.hpp:135-144)."""
        self._check(nat.lib.afqmc_calculate_overlaps(self.ctx))

    def get_vbias(self):
        """base::get_vbias(SpvnT, G, vbias, transposed=true) (vbias.hpp:45-83)."""
        self._check(nat.lib.afqmc_vbias(self.ctx))
        return self.get(nat.BUF_VBIAS)

    def fields(self, x):
        """X = x + i vbias and the hybrid weight (miniafqmc_base.cpp:281-289)."""
        xa = np.ascontiguousarray(x, dtype=np.float64)
        assert xa.shape == (self.K, self.nwalk)
        self._check(nat.lib.afqmc_fields(self.ctx, _dp(xa), 0))

    def calculate_mixed_density_matrix_full(self):
        """calculate_mixed_density_matrix(W, W_data, G, false) (mixed_density_matrix.hpp:93-101): G [2M*M][nwalk]."""
        self._check(nat.lib.afqmc_mixed_density_matrix_full(self.ctx))
        return self.get(nat.BUF_GFULL)

    def get_vbias_full(self):
        """base::get_vbias(Spvn, G, vbias, false) (vbias.hpp:65-82)."""
        self._check(nat.lib.afqmc_vbias_full(self.ctx))
        return self.get(nat.BUF_VBIAS)

    def get_vHS(self):
        """base::get_vHS(Spvn, X, vHS) (vHS.hpp:43-56)."""
        self._check(nat.lib.afqmc_vhs(self.ctx))
        return self.get(nat.BUF_VHS)

    def propagate(self):
        """afqmc_sys::propagate (afqmc_sys.hpp:146-173)."""
        self._check(nat.lib.afqmc_propagate(self.ctx))

    def save_overlaps(self):
        self._check(nat.lib.afqmc_save_overlaps(self.ctx))

    def weights(self, dt, eshift):
        e = C.c_double(eshift)
        self._check(nat.lib.afqmc_weights(self.ctx, dt, C.byref(e)))
        return e.value

    def branch(self, u):
        """Comb population control (afqmc_branch): returns the parent index of
        each new walker."""
        parent = np.zeros(self.nwalk, np.int32)
        self._check(nat.lib.afqmc_branch(self.ctx, float(u), parent.ctypes.data_as(C.POINTER(C.c_int))))
        return parent

    def set_ortho_householder(self, on):
        """Always use the blocked Householder kernel for the orthogonalisation."""
        self._check(nat.lib.afqmc_set_option(self.ctx, nat.OPT_ORTHO_HOUSEHOLDER, 1.0 if on else 0.0))

    def set_fused_propagator(self, on):
        """Split-BF16 mode: the whole propagator in one kernel (1) or as eight GEMMs (0)."""
        self._check(nat.lib.afqmc_set_option(self.ctx, nat.OPT_FUSED_PROPAGATOR, 1.0 if on else 0.0))

    def set_direct_lds(self, level=2):
        """Split-BF16 mode: 2 (default, also True) both split operands staged
        by direct global-to-LDS loads through a ring of LDS stages where they
        allow it, 1 the split left operand by direct loads into two stages,
        0 (False) every operand through registers."""
        level = 2 if level is True else int(level)
        self._check(nat.lib.afqmc_set_option(self.ctx, nat.OPT_DIRECT_LDS, float(level)))

    def set_tile(self, k):
        """Force tile candidate k (1-based) of each GEMM's kind; 0 restores
        the autotuned / heuristic choice."""
        self._check(nat.lib.afqmc_set_option(self.ctx, nat.OPT_TILE, float(k)))

    def set_three_piece_steps(self, mask):
        """With the propagator in three pieces (bit 4 of `three_piece`): a
        mask of its eight products that take three pieces, in launch order
        (bit 0 P W, bits 1-6 the Horner steps U_6 .. U_1, bit 7 P U); the
        others take two pieces.  255 (the default) is all."""
        self._check(nat.lib.afqmc_set_option(self.ctx, nat.OPT_THREE_PIECE_STEPS, float(mask)))

    def set_one_piece_steps(self, mask):
        """With the propagator in three pieces (bit 4 of `three_piece`): a
        mask of the Horner steps (bits 1-6: U_6 .. U_1) whose products take
        one piece, the hi bf16 words of the split operands."""
        self._check(nat.lib.afqmc_set_option(self.ctx, nat.OPT_ONE_PIECE_STEPS, float(mask)))

    def set_exchange_chunk(self, cols):
        """Dense exchange: accumulate the exchange GEMM over chunks of `cols`
        columns, adding the chunks' energies in FP64 (0: one pass)."""
        self._check(nat.lib.afqmc_set_option(self.ctx, nat.OPT_EXCHANGE_CHUNK, float(cols)))

    def orthogonalize(self):
        """afqmc_sys::orthogonalize (afqmc_sys.hpp:175-207)."""
        self._check(nat.lib.afqmc_orthogonalize(self.ctx))

    def calculate_energy(self):
        """afqmc_sys::calculate_energy (afqmc_sys.hpp:118-133) on the current G."""
        e = C.c_double(0.0)
        self._check(nat.lib.afqmc_energy(self.ctx, C.byref(e)))
        return e.value

    def energy_terms(self):
        """[nwalk][3] one-body, Coulomb and exchange terms of the last local energy."""
        return self.get(nat.BUF_EPARTS)

    def spvnT(self):
        """Half-rotated Cholesky matrix built at upload (rotate.hpp:65-169)."""
        from .system import Csr
        nnz = C.c_longlong(0)
        self._check(nat.lib.afqmc_get_spvnT(self.ctx, C.byref(nnz), None, None, None))
        rp = np.empty(self.K + 1, np.int32)
        col = np.empty(nnz.value, np.int32)
        val = np.empty(nnz.value, np.complex128)
        self._check(nat.lib.afqmc_get_spvnT(self.ctx, C.byref(nnz), rp.ctypes.data_as(C.POINTER(C.c_int)),
                                            col.ctypes.data_as(C.POINTER(C.c_int)), _dp(val)))
        return Csr(self.K, 2 * self.N * self.M, rp, col, val)

    # -- whole loop --------------------------------------------------------------
    def stage_fields(self, x_all):
        xa = np.ascontiguousarray(x_all, dtype=np.float64)
        assert xa.ndim == 3 and xa.shape[1:] == (self.K, self.nwalk)
        self._check(nat.lib.afqmc_stage_fields(self.ctx, _dp(xa), xa.shape[0]))

    def device_fields(self, seed, enable=True):
        """Draw the fields on the device inside each substep (the reference's
        generate_normal at miniafqmc_base.cpp:282), bit for bit the draws of
        ``system.fields(seed, K, nwalk, substep)``; no staged block needed."""
        self._check(nat.lib.afqmc_device_fields(self.ctx, int(seed), 1 if enable else 0))

    def substep(self, s, do_ortho, dt):
        self._check(nat.lib.afqmc_substep(self.ctx, s, int(do_ortho), dt))

    def eshift(self):
        e = C.c_double(0)
        self._check(nat.lib.afqmc_get_eshift(self.ctx, C.byref(e)))
        return e.value

    def set_eshift(self, e):
        self._check(nat.lib.afqmc_set_eshift(self.ctx, float(e)))

    def run(self, nsteps, nsubsteps, northo, dt, first=0, energy=True):
        """The driver's step loop (miniafqmc_base.cpp:248-350); returns Eav per step."""
        eav = np.zeros(max(nsteps, 1), np.float64)
        self._check(nat.lib.afqmc_run(self.ctx, first, nsteps, nsubsteps, northo, dt,
                                      _dp(eav) if energy else None))
        return eav[:nsteps]

    def use_graphs(self, enable=True):
        """Replay each substep from a captured HIP graph (with and without orthogonalisation)."""
        self._check(nat.lib.afqmc_use_graphs(self.ctx, 1 if enable else 0))

    def tune(self, dt):
        """Time every GEMM tile candidate once per call site; re-initialise walkers afterwards."""
        self._check(nat.lib.afqmc_tune(self.ctx, dt))

    def profile(self, on=True):
        self._check(nat.lib.afqmc_profile(self.ctx, int(on)))

    def kernel_stats(self):
        cap = 128
        names = (C.c_char * 64 * cap)()
        n = (C.c_longlong * cap)()
        ms = (C.c_double * cap)()
        cnt = C.c_int(0)
        self._check(nat.lib.afqmc_kernel_stats(self.ctx, cap, C.cast(names, C.c_void_p), n, ms, C.byref(cnt)))
        return {names[i].value.decode(): (int(n[i]), float(ms[i])) for i in range(min(cnt.value, cap))}


class AfqmcGroup:
    """Several AfqmcSys (independent systems, any sizes) advanced together:
    one HIP graph launch per substep (afqmc_group_substep)."""

    def __init__(self, systems):
        self.members = list(systems)
        arr = (C.c_void_p * len(self.members))(*[g.ctx.value for g in self.members])
        h = C.c_void_p()
        nat.check(nat.lib.afqmc_group_create(arr, len(self.members), C.byref(h)))
        self.h = h

    def _check(self, code):
        if code != 0:
            raise nat.AfqmcError(code, nat.lib.afqmc_group_error(self.h).decode())

    def substep(self, s, do_ortho, dt):
        self._check(nat.lib.afqmc_group_substep(self.h, int(s), 1 if do_ortho else 0, float(dt)))

    def sync(self):
        self._check(nat.lib.afqmc_group_sync(self.h))

    def close(self):
        if self.h:
            nat.lib.afqmc_group_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
