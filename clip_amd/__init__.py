"""clip_amd — MI355X-native coupled two-phase LBM step (Allen–Cahn + WMRT).

Python host layer over libclip_amd.so (the C-ABI of include/clip_amd.h).  It
mirrors the reference's host surface:

    InputData  (include/input/InputData.cuh)  -> Config.from_file("config.txt")
    NSAllen    (include/equation/NsAllen.cuh) -> Solver: initialCondition +
               deviceInitializer = init_fresh(), solve() = step(1),
               macroscopic fields through get()
    Reporter::sum (src/utils/Reporter.cu:95-114) -> Solver.sum("phi")
    CheckPointer::save / load (src/io/CheckPointer.cu:53-102)
                                              -> Solver.checkpoint_save / load

All compute runs in the HIP library; there is no CPU fallback.
"""
import ctypes

import numpy as np

from . import _abi
from ._abi import ClipError, CASE, SIDES, BC, GEOM, PRECISION, FIELD

__all__ = ["Config", "Solver", "InputData", "NSAllen", "ClipError", "load_library"]


def load_library():
    return _abi.load()


class Config:
    """CLIP configuration (SimParams + boundary map + geometry pool)."""

    def __init__(self, raw=None):
        self.raw = raw if raw is not None else _abi.ClipConfig()

    @classmethod
    def from_file(cls, path, dim=0):
        c = _abi.ClipConfig()
        _abi.check(_abi.load().clip_config_from_file(str(path).encode(), int(dim), ctypes.byref(c)))
        return cls(c)

    @classmethod
    def from_string(cls, text, dim=0):
        c = _abi.ClipConfig()
        _abi.check(_abi.load().clip_config_from_string(text.encode(), int(dim), ctypes.byref(c)))
        return cls(c)

    def derive(self):
        _abi.check(_abi.load().clip_config_derive(ctypes.byref(self.raw)))
        return self

    def copy(self):
        c = _abi.ClipConfig()
        ctypes.memmove(ctypes.byref(c), ctypes.byref(self.raw), ctypes.sizeof(c))
        return Config(c)

    @property
    def dim(self):
        return self.raw.dim

    @property
    def q(self):
        return 9 if self.raw.dim == 2 else 19

    def extents(self):
        N = list(self.raw.N)
        if self.raw.dim == 2:
            return (int(N[0]) + 2, int(N[1]) + 2, 1)
        return (int(N[0]), int(N[1]), int(N[2]))

    def interior_slices(self):
        ex = self.extents()
        if self.raw.dim == 2:
            return (slice(1, ex[0] - 1), slice(1, ex[1] - 1), slice(0, 1))
        return tuple(slice(1, e - 1) for e in ex)

    def __getattr__(self, name):
        if name == "raw":
            raise AttributeError(name)
        return getattr(self.raw, name)

    def with_params(self, sigma=None, mobility=None):
        """Copy with the surface tension and/or the mobility replaced
        (derived beta, kappa follow: InputData.cu:76-114).  Both are kept as
        explicit overrides, so a case that derives them (RTI derives the
        mobility from Pe) uses the values given."""
        c = self.copy()
        if sigma is not None:
            c.raw.has_sigma = 1
            c.raw.sigma_override = float(sigma)
        if mobility is not None:
            c.raw.has_mobility_override = 1
            c.raw.mobility_override = float(mobility)
        return c.derive()

    def sdf(self, x, y, z=0.0, geometry_id=0):
        """Geometry::sdf (Geometry.cuh:102-161): signed distance to the first
        geometry object with this id (positive outside; 1e10 if none)."""
        out = ctypes.c_double()
        _abi.check(_abi.load().clip_geometry_sdf(ctypes.byref(self.raw), int(geometry_id), float(x), float(y),
                                                 float(z), ctypes.byref(out)))
        return out.value

    def initial_phi(self):
        """NSAllen::initialCondition (NsAllen.cu:768-824): the initial phase
        field on every node, shape (ext_x, ext_y, ext_z, 1)."""
        ex = self.extents()
        out = np.empty(ex[0] * ex[1] * ex[2], dtype=np.float64)
        _abi.check(_abi.load().clip_initial_phi(ctypes.byref(self.raw),
                                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out.reshape(ex[0], ex[1], ex[2], 1)

    def interior_shape(self):
        ex = self.extents()
        return tuple(e - 2 for e in ex[: self.raw.dim])

    def set(self, **kw):
        for k, v in kw.items():
            cur = getattr(self.raw, k)
            if isinstance(cur, ctypes.Array):
                for i, x in enumerate(v):
                    cur[i] = x
            else:
                setattr(self.raw, k, v)
        return self


InputData = Config


def _ptr(addr):
    return None if addr is None else ctypes.c_void_p(int(addr))


class Solver:
    """One solver handle (device buffers for one configuration)."""

    def __init__(self, config, precision="f64"):
        self.config = config
        self.precision = precision
        self._lib = _abi.load()
        h = ctypes.c_void_p()
        _abi.check(self._lib.clip_create(ctypes.byref(config.raw), PRECISION[precision], ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.clip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # NSAllen::initialCondition + deviceInitializer
    def init_fresh(self):
        _abi.check(self._lib.clip_init_fresh(self._h))
        return self

    def init_from_phi(self, phi):
        """NSAllen::deviceInitializer from a phase field on every node
        (shape (ext_x, ext_y, ext_z[, 1]), CLIP's order, ghosts included)."""
        ex = self.config.extents()
        arr = np.ascontiguousarray(phi, dtype=np.float64).ravel()
        if arr.size != ex[0] * ex[1] * ex[2]:
            raise ValueError(f"phi must hold {ex[0] * ex[1] * ex[2]} nodes, got {arr.size}")
        _abi.check(self._lib.clip_init_from_phi(self._h, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return self

    # Solver's boundary launchers (Solver.cuh:61-107) on device arrays the
    # caller owns: FP64, CLIP's AoS layout, ghosts included; arguments are
    # device addresses (int) or None.
    def periodic_boundary(self, ndof, dev_a, dev_b=None):
        _abi.check(self._lib.clip_periodic_boundary(self._h, int(ndof), _ptr(dev_a), _ptr(dev_b)))

    def wall_boundary(self, dev_a, dev_a_post, dev_b=None, dev_b_post=None):
        _abi.check(self._lib.clip_wall_boundary(self._h, _ptr(dev_a), _ptr(dev_a_post), _ptr(dev_b),
                                                _ptr(dev_b_post)))

    def slip_wall_boundary(self, dev_a, dev_a_post, dev_b=None, dev_b_post=None):
        _abi.check(self._lib.clip_slip_wall_boundary(self._h, _ptr(dev_a), _ptr(dev_a_post), _ptr(dev_b),
                                                     _ptr(dev_b_post)))

    def free_convect_boundary(self, dev_vel, dev_a, dev_a_prev, dev_b=None, dev_b_prev=None):
        _abi.check(self._lib.clip_free_convect_boundary(self._h, _ptr(dev_vel), _ptr(dev_a), _ptr(dev_a_prev),
                                                        _ptr(dev_b), _ptr(dev_b_prev)))

    def neumann_boundary(self, dev_a, dev_b=None):
        _abi.check(self._lib.clip_neumann_boundary(self._h, _ptr(dev_a), _ptr(dev_b)))

    def velocity_boundary(self, dev_c, dev_f, dev_g):
        _abi.check(self._lib.clip_velocity_boundary(self._h, _ptr(dev_c), _ptr(dev_f), _ptr(dev_g)))

    def mirror_boundary(self, dev_a):
        _abi.check(self._lib.clip_mirror_boundary(self._h, _ptr(dev_a)))

    # nsteps x NSAllen::solve
    def step(self, nsteps=1):
        _abi.check(self._lib.clip_step(self._h, int(nsteps)))
        return self

    solve = step

    def synchronize(self):
        _abi.check(self._lib.clip_synchronize(self._h))

    def set_stream(self, stream_ptr):
        """Bind the handle to a hipStream_t of the HIP runtime this library
        links (/opt/rocm).  A PyTorch wheel that bundles its own libamdhip64
        has its own streams, which are not valid here."""
        _abi.check(self._lib.clip_set_stream(self._h, ctypes.c_void_p(int(stream_ptr))))

    def info(self):
        inf = _abi.ClipInfo()
        _abi.check(self._lib.clip_get_info(self._h, ctypes.byref(inf)))
        return inf

    def ndof(self, field):
        d = self.config.dim
        return {"phi": 1, "rho": 1, "p": 1, "mu": 1, "u": d, "dphi": d, "normal": d}.get(field, self.config.q)

    def get(self, field, layout="grid"):
        """Field as float64.  layout="grid": array (ext_x, ext_y, ext_z, ndof)
        (CLIP's AoS order); "aos": flat CLIP AoS; "soa": flat dof-major."""
        ex = self.config.extents()
        nd = self.ndof(field)
        out = np.empty(ex[0] * ex[1] * ex[2] * nd, dtype=np.float64)
        lay = 1 if layout == "soa" else 0
        _abi.check(self._lib.clip_get_field(self._h, FIELD[field], lay,
                                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        if layout == "grid":
            return out.reshape(ex[0], ex[1], ex[2], nd)
        return out

    def set(self, field, values, layout="grid"):
        """Overwrites one of the reference's device arrays (clip_set_field):
        phi, u, p, f, g, f_post or g_post, ghost cells included, in the same
        shape get() returns.  The next step collides from them; call
        macroscopic() to recompute phi, p and u from f_post / g_post."""
        if field not in ("phi", "u", "p", "f", "g", "f_post", "g_post"):
            raise ValueError(f"{field!r} cannot be set (rho, mu, dphi and normal are derived from phi)")
        arr = np.ascontiguousarray(values, dtype=np.float64).ravel()
        ex = self.config.extents()
        want = ex[0] * ex[1] * ex[2] * self.ndof(field)
        if arr.size != want:
            raise ValueError(f"{field}: {arr.size} values given, the grid holds {want} (ghost cells included)")
        lay = 1 if layout == "soa" else 0
        _abi.check(self._lib.clip_set_field(self._h, FIELD[field], lay,
                                            arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def macroscopic(self):
        """NSAllen::macroscopic: phi, mu, grad phi, p and u recomputed from the
        current f_post / g_post (u_old = the current u)."""
        _abi.check(self._lib.clip_macroscopic(self._h))
        return self

    def interior(self, field):
        a = self.get(field)
        return a[self.config.interior_slices()]

    def sum(self, field="phi"):
        out = ctypes.c_double()
        _abi.check(self._lib.clip_sum(self._h, FIELD[field], ctypes.byref(out)))
        return out.value

    def check_health(self):
        """Reporter::print's divergence check; raises ClipError(ERR_DIVERGED)."""
        out = ctypes.c_double()
        _abi.check(self._lib.clip_check_health(self._h, ctypes.byref(out)))
        return out.value

    def write_vts(self, path):
        """VTSwriter: Velocity, C, P, Rho of the interior nodes as ASCII VTK."""
        _abi.check(self._lib.clip_write_vts(self._h, str(path).encode()))

    def checkpoint_save(self, directory):
        _abi.check(self._lib.clip_checkpoint_save(self._h, str(directory).encode()))

    def checkpoint_load(self, directory):
        _abi.check(self._lib.clip_checkpoint_load(self._h, str(directory).encode()))

    def enable_graph(self, on=True):
        """Step through a captured HIP graph of two steps (clip_enable_graph)."""
        _abi.check(self._lib.clip_enable_graph(self._h, 1 if on else 0))
        return self

    def enable_timing(self, on=True):
        _abi.check(self._lib.clip_enable_timing(self._h, 1 if on else 0))

    # ---- reverse mode (fully periodic domains, "f64" / "f32" handles) ----
    def adjoint_forward(self, phi0, nsteps, checkpoint_every=50, inner_every=0):
        """Runs nsteps steps from the interior phase field phi0 (array of the
        interior shape, index [i, j(, k)]) and returns phi after the last
        step (same shape); keeps states for adjoint_backward (every
        checkpoint_every steps, and with 0 < inner_every < checkpoint_every a
        second level inside each recomputed segment)."""
        _abi.check(self._lib.clip_adjoint_set_nesting(self._h, int(inner_every)))
        a = self._interior_array(phi0)
        out = np.empty_like(a)
        _abi.check(self._lib.clip_adjoint_forward(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                  int(nsteps), int(checkpoint_every),
                                                  out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                  _abi.INTERIOR_X_SLOWEST))
        return out

    def adjoint_backward(self, grad_phi, normal_floor=0.0):
        """dL/dphi0, dL/dsigma, dL/dmobility for dL/dphi(last step) = grad_phi.
        normal_floor = 0: exact derivative of the steps; > 0 (e.g. 0.1):
        bounded derivative of the interface normal on flat plateaus
        (clip_adjoint_set_normal_floor)."""
        _abi.check(self._lib.clip_adjoint_set_normal_floor(self._h, float(normal_floor)))
        g = self._interior_array(grad_phi)
        out = np.empty_like(g)
        gs, gm = ctypes.c_double(), ctypes.c_double()
        _abi.check(self._lib.clip_adjoint_backward(self._h, g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                   ctypes.byref(gs), ctypes.byref(gm), _abi.INTERIOR_X_SLOWEST))
        return out, gs.value, gm.value

    def _interior_array(self, x):
        """x as a C-ordered float64 array of the interior shape (index [i, j(,
        k)]: the library's x-slowest interior order, no copy when x already
        is one)."""
        return np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(self.config.interior_shape()))

    def adjoint_inlet_gradient(self):
        """dL/dA_m of the inlet modes (inletModeAmplitude) from the last
        adjoint_backward (numpy array of length inlet_modes)."""
        n = int(self.config.raw.inlet_modes)
        out = np.zeros(max(n, 1))
        _abi.check(self._lib.clip_adjoint_inlet_gradient(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                         n))
        return out[:n]

    def last_timing(self):
        vals = [ctypes.c_double() for _ in range(4)]
        _abi.check(self._lib.clip_last_step_timing(self._h, *[ctypes.byref(v) for v in vals]))
        return {"total_ms": vals[0].value, "ghost_ms": vals[1].value,
                "phase_ms": vals[2].value, "fused_ms": vals[3].value}


NSAllen = Solver
