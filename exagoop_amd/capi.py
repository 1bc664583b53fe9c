"""ctypes binding of the exagoop_amd C ABI (include/exagoop_amd.h).

Loads the in-tree libexagoop_amd.so.  There is no fallback: a missing or
unloadable library raises ImportError, and every non-zero status code raises
MPMError with the library's error string.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libexagoop_amd.so")

NREAL = 46
NINT = 3

# Particle real fields (mpm_specs.H:17-73)
R_RADIUS, R_VEL, R_VELP, R_STRAINRATE, R_STRAIN, R_STRESS, R_DEFGRAD = 0, 1, 4, 7, 13, 19, 25
R_VOLUME, R_MASS, R_DENSITY, R_JACOBIAN, R_PRESSURE, R_VOLINIT = 34, 35, 36, 37, 38, 39
R_E, R_NU, R_BULK, R_GAMMA, R_VISC, R_YACC = 40, 41, 42, 43, 44, 45
I_PHASE, I_RIGID_ID, I_MODEL = 0, 1, 2

# Nodal components (constants.H:36-55)
N_MASS, N_VEL, N_FRC, N_DVEL, N_MASS_OLD = 0, 1, 4, 7, 10

BC_PERIODIC, BC_NOSLIP, BC_SLIP, BC_PARTIALSLIP, BC_OUTFLOW = 0, 1, 2, 3, 4

# Thermal module (the USE_TEMP build): particle fields (realData 46..52,
# mpm_specs.H:52-58), nodal components (constants.H:58-63), wall conditions
# (constants.H:95-99).
NTHERMAL = 7
TH_TEMPERATURE, TH_SPECIFIC_HEAT, TH_CONDUCTIVITY, TH_HEAT_FLUX, TH_HEAT_SOURCE = 0, 1, 2, 3, 6
N_MASS_SPHEAT, N_MASS_SPHEAT_TEMP, N_TEMPERATURE = 20, 21, 22
N_SOURCE_TEMP, N_DELTA_TEMPERATURE, N_MASS_CONDUCTIVITY = 23, 24, 25
BC_TEMP_PERIODIC, BC_TEMP_ISOTHERMAL, BC_TEMP_ADIABATIC = 0, 1, 2
BC_TEMP_USERFLUX, BC_TEMP_CONVECTION = 3, 4
LS_TEMP_ADIABATIC, LS_TEMP_ISOTHERMAL, LS_TEMP_HEATFLUX, LS_TEMP_CONVECTIVE = 0, 1, 3, 4
USL, MUSL = 0, 1

OPT_PROFILE, OPT_KEEP_FORCES, OPT_SORTED, OPT_FAST, OPT_FUSED_DT = 1, 2, 3, 4, 5
OPT_DETERMINISTIC = 6
OPT_COARSE_BLEND = 7
OPT_TILE_WAVES = 8
LS_SPHERE, LS_PLANE, LS_CYLINDER, LS_BOX = 0, 1, 2, 3
MLP_NPARAMS = 11331
MLP_IN = 128

STATUS = {0: "OK", -1: "E_ARG", -2: "E_BC", -3: "E_ORDER", -4: "E_MODEL", -5: "E_OOM",
          -6: "E_HIP", -7: "E_NAN", -8: "E_UNSUPPORTED", -9: "E_NODEVICE"}

_d3 = ctypes.c_double * 3
_i3 = ctypes.c_int * 3
_d9 = ctypes.c_double * 9


class Specs(ctypes.Structure):
    """mpm_specs (include/exagoop_amd.h); mirrors the MPMspecs step subset."""
    _fields_ = [
        ("plo", _d3), ("phi", _d3), ("ncells", _i3), ("periodic", _i3),
        ("order_scheme", ctypes.c_int), ("num_redist", ctypes.c_int),
        ("gravity", _d3), ("external_loads", ctypes.c_int),
        ("force_slab_lo", _d3), ("force_slab_hi", _d3), ("extforce", _d3),
        ("mass_tolerance", ctypes.c_double), ("alpha_pic_flip", ctypes.c_double),
        ("stress_update_scheme", ctypes.c_int), ("fixed_timestep", ctypes.c_int),
        ("timestep", ctypes.c_double), ("CFL", ctypes.c_double),
        ("dt_min_limit", ctypes.c_double), ("dt_max_limit", ctypes.c_double),
        ("applied_strainrate", ctypes.c_double), ("applied_strainrate_time", ctypes.c_double),
        ("calculate_strain_based_on_delta", ctypes.c_int),
        ("bclo", _i3), ("bchi", _i3), ("wall_mu_lo", _d3), ("wall_mu_hi", _d3),
        ("wall_vel_lo", _d9), ("wall_vel_hi", _d9),
    ]

    def copy(self):
        s = Specs()
        ctypes.pointer(s)[0] = self
        return s


class ThermalSpecs(ctypes.Structure):
    """mpm_thermal_specs: per-face temperature conditions and parameters."""
    _fields_ = [
        ("bclo", _i3), ("bchi", _i3), ("T_wall_lo", _d3), ("T_wall_hi", _d3),
        ("flux_lo", _d3), ("flux_hi", _d3), ("h_lo", _d3), ("h_hi", _d3),
        ("T_inf_lo", _d3), ("T_inf_hi", _d3),
    ]


def default_thermal_specs(specs):
    """Reference defaults (mpm_specs.H:227-234): periodic faces periodic,
    the others adiabatic, all parameters zero."""
    t = ThermalSpecs()
    for d in range(3):
        code = BC_TEMP_PERIODIC if specs.periodic[d] else BC_TEMP_ADIABATIC
        t.bclo[d] = code
        t.bchi[d] = code
    return t


def default_specs():
    """Reference MPMspecs defaults (mpm_specs.H:78-220), built in Python so the
    oracle can be driven without the HIP library."""
    s = Specs()
    s.gravity[:] = [0.0, -9.81, 0.0]
    s.num_redist = 10
    s.order_scheme = 1
    s.mass_tolerance = 1e-8
    s.alpha_pic_flip = 0.95
    s.stress_update_scheme = USL
    s.CFL = 0.1
    s.dt_min_limit = 1e-12
    s.dt_max_limit = 1.0
    for d in range(3):
        s.bclo[d] = BC_SLIP
        s.bchi[d] = BC_SLIP
        s.wall_mu_lo[d] = 0.1
        s.wall_mu_hi[d] = 0.1
    return s


UDF_FN = ctypes.CFUNCTYPE(None, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                          ctypes.c_double, ctypes.POINTER(ctypes.c_double))

# Exported symbols and their signatures: name -> (restype, argtypes)
_P = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)
_LP = ctypes.POINTER(ctypes.c_int64)
_FP = ctypes.POINTER(ctypes.c_float)
SIGNATURES = {
    "mpm_specs_default": (None, [ctypes.POINTER(Specs)]),
    "mpm_order_directional": (ctypes.c_int, [ctypes.POINTER(Specs), _IP]),
    "mpm_create": (ctypes.c_int, [ctypes.POINTER(Specs), ctypes.POINTER(_P)]),
    "mpm_destroy": (None, [_P]),
    "mpm_upload_particles": (ctypes.c_int, [_P, ctypes.c_int64, _DP, _DP, _IP]),
    "mpm_num_particles": (ctypes.c_int, [_P, _LP]),
    "mpm_download_particles": (ctypes.c_int, [_P, _DP, _DP, _IP, _LP]),
    "mpm_get_nodal": (ctypes.c_int, [_P, ctypes.c_int, _DP]),
    "mpm_set_nodal": (ctypes.c_int, [_P, ctypes.c_int, _DP]),
    "mpm_read_particle_file": (ctypes.c_int, [ctypes.c_char_p, _LP, _DP, _DP, _IP, _DP]),
    "mpm_load_particle_file": (ctypes.c_int, [_P, ctypes.c_char_p, _DP]),
    "mpm_autogen_particles": (ctypes.c_int, [ctypes.POINTER(Specs), _DP, _DP, _DP, _IP,
                                             ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_int, _LP, _DP, _DP, _IP,
                                             _DP]),
    "mpm_write_particle_file": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int64, _DP, _DP, _IP]),
    "mpm_write_checkpoint": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_int]),
    "mpm_levelset_analytic": (ctypes.c_int, [ctypes.POINTER(Specs), ctypes.c_int, _DP,
                                             ctypes.c_int, _DP]),
    "mpm_remove_particles_inside_levelset": (ctypes.c_int, [_P, _LP]),
    "mpm_read_checkpoint": (ctypes.c_int, [_P, ctypes.c_char_p, _IP]),
    "mpm_write_ascii_particles": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_double,
                                                 ctypes.c_char_p, ctypes.c_int64]),
    "mpm_get_particle_field": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int64, _DP]),
    "mpm_reset_nodal": (ctypes.c_int, [_P]),
    "mpm_p2g": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "mpm_backup_velocity": (ctypes.c_int, [_P]),
    "mpm_nodal_update": (ctypes.c_int, [_P, ctypes.c_double]),
    "mpm_apply_nodal_bcs": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_double]),
    "mpm_g2p": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_double]),
    "mpm_move": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_int]),
    "mpm_redistribute": (ctypes.c_int, [_P, ctypes.c_int]),
    "mpm_update_volume": (ctypes.c_int, [_P]),
    "mpm_stress": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_double]),
    "mpm_calculate_dt": (ctypes.c_int, [_P, _DP]),
    "mpm_init_internal_forces": (ctypes.c_int, [_P]),
    "mpm_step": (ctypes.c_int, [_P, ctypes.c_int]),
    "mpm_get_time": (ctypes.c_int, [_P, _DP, _LP, _DP]),
    "mpm_add_levelset_body": (ctypes.c_int, [_P, _DP, ctypes.c_int, ctypes.c_double, _DP]),
    "mpm_set_wall_udf": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, UDF_FN]),
    "mpm_energies": (ctypes.c_int, [_P, _DP]),
    "mpm_energies_noalloc": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "mpm_set_option": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int]),
    "mpm_last_step_timing": (ctypes.c_int, [_P, _DP, _DP]),
    "mpm_last_error": (ctypes.c_char_p, []),
    "mpm_mlp_create": (ctypes.c_int, [ctypes.POINTER(_P)]),
    "mpm_mlp_destroy": (None, [_P]),
    "mpm_mlp_set_params": (ctypes.c_int, [_P, _FP, _DP]),
    "mpm_mlp_get_params": (ctypes.c_int, [_P, _FP, _DP]),
    "mpm_mlp_forward": (ctypes.c_int, [_P, ctypes.c_int64, _FP, _FP]),
    "mpm_mlp_gradient": (ctypes.c_int, [_P, ctypes.c_int64, _FP, _FP, _FP, _DP]),
    "mpm_mlp_adam": (ctypes.c_int, [_P, _FP, ctypes.c_float, ctypes.c_int]),
    "mpm_dataset_create": (ctypes.c_int, [ctypes.c_int64, ctypes.POINTER(_P)]),
    "mpm_dataset_destroy": (None, [_P]),
    "mpm_dataset_size": (ctypes.c_int, [_P, _LP]),
    "mpm_dataset_add": (ctypes.c_int, [_P, ctypes.c_int64, _FP, _FP]),
    "mpm_dataset_target_rms": (ctypes.c_int, [_P, _DP]),
    "mpm_learn_features": (ctypes.c_int, [_P, _P, ctypes.c_int64, _LP, _FP]),
    "mpm_learn_collect": (ctypes.c_int, [_P, _P, _P, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_uint64]),
    "mpm_mlp_train": (ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                     ctypes.c_uint64, _DP]),
    "mpm_step_learned": (ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.c_int, _DP, _IP]),
    "mpm_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(Specs), _LP,
                                        ctypes.POINTER(_P)]),
    "mpm_batch_destroy": (None, [_P]),
    "mpm_batch_upload": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int64, _DP, _DP, _IP]),
    "mpm_batch_init": (ctypes.c_int, [_P]),
    "mpm_batch_step": (ctypes.c_int, [_P, ctypes.c_int]),
    "mpm_batch_num_particles": (ctypes.c_int, [_P, ctypes.c_int, _LP]),
    "mpm_batch_download": (ctypes.c_int, [_P, ctypes.c_int, _DP, _DP, _IP, _LP]),
    "mpm_batch_get_nodal": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, _DP]),
    "mpm_batch_get_time": (ctypes.c_int, [_P, ctypes.c_int, _DP, _LP, _DP]),
    "mpm_batch_get_tuning": (ctypes.c_int, [_P, ctypes.c_void_p]),
    "mpm_batch_set_tuning": (ctypes.c_int, [_P, ctypes.c_void_p]),
    "mpm_batch_autotune": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_void_p, _DP]),
    "mpm_batch_stats": (ctypes.c_int, [_P, _DP, _LP, _IP]),
    "mpm_batch_profile": (ctypes.c_int, [_P, ctypes.c_int, _DP, _DP, _LP, _DP, _DP]),
    "mpm_batch_set_option": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int]),
    "mpm_thermal_specs_default": (None, [ctypes.POINTER(Specs), ctypes.POINTER(ThermalSpecs)]),
    "mpm_enable_thermal": (ctypes.c_int, [_P, ctypes.POINTER(ThermalSpecs)]),
    "mpm_upload_thermal": (ctypes.c_int, [_P, ctypes.c_int64, _DP]),
    "mpm_download_thermal": (ctypes.c_int, [_P, ctypes.c_int64, _DP]),
    "mpm_set_wall_temp_udf": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, UDF_FN]),
    "mpm_set_levelset_body_temp": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                  ctypes.c_double, ctypes.c_double,
                                                  ctypes.c_double]),
    "mpm_read_particle_file_thermal": (ctypes.c_int, [ctypes.c_char_p, _LP, _DP, _DP, _IP, _DP,
                                                      _DP]),
    "mpm_load_particle_file_thermal": (ctypes.c_int, [_P, ctypes.c_char_p, _DP]),
    "mpm_write_particle_file_thermal": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int64, _DP, _DP,
                                                       _IP, _DP]),
}


class MPMError(RuntimeError):
    pass


_lib = None


def load():
    """Load libexagoop_amd.so once; raise ImportError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    try:
        lib = ctypes.CDLL(LIB_PATH)
    except OSError as e:
        raise ImportError(f"cannot load {LIB_PATH}: {e}") from e
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def dptr(a):
    return a.ctypes.data_as(_DP)


def iptr(a):
    return a.ctypes.data_as(_IP)


def lptr(a):
    return a.ctypes.data_as(_LP)


class _Base:
    """Shared wrapper logic for the product context and the oracle context."""

    prefix = "mpm_"

    def _call(self, name, *args):
        rc = getattr(self._lib, self.prefix + name)(*args)
        if rc != 0:
            raise MPMError(f"{self.prefix}{name} -> {STATUS.get(rc, rc)}: {self._errstr()}")
        return rc

    def _errstr(self):
        return ""

    def close(self):
        if getattr(self, "_ctx", None):
            getattr(self._lib, self.prefix + "destroy")(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, pos, rdata, idata):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        rdata = np.ascontiguousarray(rdata, dtype=np.float64)
        idata = np.ascontiguousarray(idata, dtype=np.int32)
        np_ = pos.shape[1]
        assert pos.shape == (3, np_) and rdata.shape == (NREAL, np_) and idata.shape == (NINT, np_)
        self._call("upload_particles", self._ctx, np_, dptr(pos), dptr(rdata), iptr(idata))

    def num_particles(self):
        n = ctypes.c_int64(0)
        self._call("num_particles", self._ctx, ctypes.byref(n))
        return n.value

    def download(self):
        n = self.num_particles()
        pos = np.zeros((3, n))
        rdata = np.zeros((NREAL, n))
        idata = np.zeros((NINT, n), dtype=np.int32)
        ids = np.zeros(n, dtype=np.int64)
        if n:
            self._call("download_particles", self._ctx, dptr(pos), dptr(rdata), iptr(idata),
                       lptr(ids))
        return pos, rdata, idata, ids

    def nodal(self, comp):
        out = np.zeros(self.nn)
        self._call("get_nodal", self._ctx, comp, dptr(out))
        return out

    def step(self, n=1):
        self._call("step", self._ctx, n)

    def init_internal_forces(self):
        self._call("init_internal_forces", self._ctx)

    def reset_nodal(self):
        self._call("reset_nodal", self._ctx)

    def p2g(self, um, uv, uf):
        self._call("p2g", self._ctx, um, uv, uf)

    def backup_velocity(self):
        self._call("backup_velocity", self._ctx)

    def nodal_update(self, dt):
        self._call("nodal_update", self._ctx, dt)

    def apply_nodal_bcs(self, dt, t):
        self._call("apply_nodal_bcs", self._ctx, dt, t)

    def g2p(self, uv, us, dt):
        self._call("g2p", self._ctx, uv, us, dt)

    def move(self, dt, step):
        self._call("move", self._ctx, dt, step)

    def redistribute(self, step):
        self._call("redistribute", self._ctx, step)

    def update_volume(self):
        self._call("update_volume", self._ctx)

    def stress(self, dt, time):
        self._call("stress", self._ctx, dt, time)

    def calculate_dt(self):
        v = ctypes.c_double(0)
        self._call("calculate_dt", self._ctx, ctypes.byref(v))
        return v.value

    def time(self):
        t = ctypes.c_double(0)
        s = ctypes.c_int64(0)
        dt = ctypes.c_double(0)
        self._call("get_time", self._ctx, ctypes.byref(t), ctypes.byref(s), ctypes.byref(dt))
        return t.value, s.value, dt.value

    def add_levelset_body(self, phi, bc, mu, wall_vel=(0.0, 0.0, 0.0)):
        phi = np.ascontiguousarray(phi, dtype=np.float64).ravel()
        assert phi.size == self.nn
        w = np.asarray(wall_vel, dtype=np.float64)
        self._call("add_levelset_body", self._ctx, dptr(phi), bc, mu, dptr(w))

    def set_wall_udf(self, d, hi, fn):
        """fn(x, y, z, t) -> (vx, vy, vz); kept alive by this object."""
        def _cb(x, y, z, t, out):
            v = fn(x, y, z, t)
            out[0], out[1], out[2] = v[0], v[1], v[2]
        cb = UDF_FN(_cb)
        self._udf_refs = getattr(self, "_udf_refs", [])
        self._udf_refs.append(cb)
        self._call("set_wall_udf", self._ctx, d, hi, cb)

    def set_wall_temp_udf(self, d, hi, fn):
        """fn(x, y, z, t) -> (v0, v1): T_wall, flux or h; T_inf (convection)."""
        def _cb(x, y, z, t, out):
            v = fn(x, y, z, t)
            out[0], out[1] = v[0], v[1]
        cb = UDF_FN(_cb)
        self._udf_refs = getattr(self, "_udf_refs", [])
        self._udf_refs.append(cb)
        self._call("set_wall_temp_udf", self._ctx, d, hi, cb)

    def set_levelset_body_temp(self, body, temp_bc, T_wall=0.0, heat_flux=0.0, h_conv=0.0,
                               T_inf=0.0):
        self._call("set_levelset_body_temp", self._ctx, body, temp_bc, T_wall, heat_flux, h_conv,
                   T_inf)

    def energies(self):
        out = np.zeros(6)
        self._call("energies", self._ctx, dptr(out))
        return out

    def enable_thermal(self, tspecs=None):
        """Turn on the thermal stages; tspecs defaults to the reference's."""
        t = default_thermal_specs(self.specs) if tspecs is None else tspecs
        self._call("enable_thermal", self._ctx, ctypes.byref(t))
        self._n_thermal = None

    def upload_thermal(self, th):
        """th: (NTHERMAL, n_upload) by upload index."""
        th = np.ascontiguousarray(th, dtype=np.float64)
        assert th.ndim == 2 and th.shape[0] == NTHERMAL
        self._call("upload_thermal", self._ctx, th.shape[1], dptr(th))
        self._n_thermal = th.shape[1]

    def download_thermal(self, n_upload=None):
        """(NTHERMAL, n_upload) by upload index, NaN for removed particles."""
        n = self._n_thermal if n_upload is None else n_upload
        assert n is not None, "pass n_upload (no thermal upload yet)"
        out = np.zeros((NTHERMAL, n))
        self._call("download_thermal", self._ctx, n, dptr(out))
        return out


class Context(_Base):
    """An exagoop_amd context on the current HIP device."""

    prefix = "mpm_"

    def __init__(self, specs):
        self._lib = load()
        self.specs = specs.copy()
        ctx = ctypes.c_void_p()
        self._ctx = None
        self._call("create", ctypes.byref(specs), ctypes.byref(ctx))
        self._ctx = ctx
        n = specs.ncells
        self.nn = (n[0] + 1) * (n[1] + 1) * (n[2] + 1)

    def _errstr(self):
        return self._lib.mpm_last_error().decode()

    def set_nodal(self, comp, values):
        a = np.ascontiguousarray(values, dtype=np.float64)
        assert a.shape == (self.nn,)
        self._call("set_nodal", self._ctx, comp, dptr(a))

    def set_option(self, opt, value):
        self._call("set_option", self._ctx, opt, value)

    def write_checkpoint(self, path, output_it=0):
        self._call("write_checkpoint", self._ctx, os.fsencode(path), output_it)

    def write_ascii_particles(self, prefix, time):
        """Returns the path written."""
        buf = ctypes.create_string_buffer(4096)
        self._call("write_ascii_particles", self._ctx, os.fsencode(prefix), time, buf, 4096)
        return buf.value.decode()

    def read_checkpoint(self, path):
        """Returns the output index stored in the checkpoint."""
        oit = ctypes.c_int(0)
        self._call("read_checkpoint", self._ctx, os.fsencode(path), ctypes.byref(oit))
        return oit.value

    def remove_particles_inside_levelset(self):
        n = ctypes.c_int64(0)
        self._call("remove_particles_inside_levelset", self._ctx, ctypes.byref(n))
        return n.value

    def load_particle_file(self, path, thermal=False):
        """mpm_load_particle_file (thermal: the USE_TEMP layout, also
        uploading the thermal fields); returns (total mass, total volume)."""
        tot = np.zeros(2)
        name = "load_particle_file_thermal" if thermal else "load_particle_file"
        self._call(name, self._ctx, os.fsencode(path), dptr(tot))
        if thermal:
            self._n_thermal = self.num_particles()
        return float(tot[0]), float(tot[1])

    def particle_field(self, field, n_upload):
        """Real field `field` of every particle by upload index (NaN: removed)."""
        out = np.empty(n_upload)
        self._call("get_particle_field", self._ctx, field, n_upload, dptr(out))
        return out

    def last_step_timing(self):
        tot = ctypes.c_double(0)
        k = np.zeros(4)
        self._call("last_step_timing", self._ctx, ctypes.byref(tot), dptr(k))
        return tot.value, k


def fptr(a):
    return a.ctypes.data_as(_FP)


def _check(lib, rc, what):
    if rc != 0:
        raise MPMError(f"{what} -> {STATUS.get(rc, rc)}: {lib.mpm_last_error().decode()}")


class MLP:
    """The learned correction's network on the device."""

    def __init__(self, params=None, scales=(1.0, 1.0, 1.0)):
        self._lib = load()
        p = _P()
        _check(self._lib, self._lib.mpm_mlp_create(ctypes.byref(p)), "mpm_mlp_create")
        self._p = p
        if params is not None:
            self.set_params(params, scales)

    def close(self):
        if getattr(self, "_p", None):
            self._lib.mpm_mlp_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params, scales):
        params = np.ascontiguousarray(params, dtype=np.float32)
        assert params.size == MLP_NPARAMS
        sc = np.ascontiguousarray(scales, dtype=np.float64)
        _check(self._lib, self._lib.mpm_mlp_set_params(self._p, fptr(params), dptr(sc)),
               "mpm_mlp_set_params")

    def get_params(self):
        params = np.zeros(MLP_NPARAMS, dtype=np.float32)
        sc = np.zeros(3)
        _check(self._lib, self._lib.mpm_mlp_get_params(self._p, fptr(params), dptr(sc)),
               "mpm_mlp_get_params")
        return params, sc

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], 3), dtype=np.float32)
        _check(self._lib, self._lib.mpm_mlp_forward(self._p, X.shape[0], fptr(X), fptr(Y)),
               "mpm_mlp_forward")
        return Y

    def gradient(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.zeros(MLP_NPARAMS, dtype=np.float32)
        loss = ctypes.c_double(0)
        _check(self._lib, self._lib.mpm_mlp_gradient(self._p, X.shape[0], fptr(X), fptr(T),
                                                     fptr(g), ctypes.byref(loss)),
               "mpm_mlp_gradient")
        return loss.value, g

    def adam(self, grad, lr, t):
        grad = np.ascontiguousarray(grad, dtype=np.float32)
        _check(self._lib, self._lib.mpm_mlp_adam(self._p, fptr(grad), lr, t), "mpm_mlp_adam")

    def train(self, dataset, epochs, batch, lr, seed=0):
        loss = np.zeros(max(epochs, 1))
        _check(self._lib, self._lib.mpm_mlp_train(self._p, dataset._p, epochs, batch, lr, seed,
                                                  dptr(loss)), "mpm_mlp_train")
        return loss[:epochs]


class Dataset:
    """Training rows on the device."""

    def __init__(self, capacity):
        self._lib = load()
        p = _P()
        _check(self._lib, self._lib.mpm_dataset_create(capacity, ctypes.byref(p)),
               "mpm_dataset_create")
        self._p = p

    def close(self):
        if getattr(self, "_p", None):
            self._lib.mpm_dataset_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def size(self):
        n = ctypes.c_int64(0)
        _check(self._lib, self._lib.mpm_dataset_size(self._p, ctypes.byref(n)),
               "mpm_dataset_size")
        return n.value

    def add(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        _check(self._lib, self._lib.mpm_dataset_add(self._p, X.shape[0], fptr(X), fptr(T)),
               "mpm_dataset_add")

    def target_rms(self):
        v = ctypes.c_double(0)
        _check(self._lib, self._lib.mpm_dataset_target_rms(self._p, ctypes.byref(v)),
               "mpm_dataset_target_rms")
        return v.value


def learn_features(ctx, mlp, nodes):
    nodes = np.ascontiguousarray(nodes, dtype=np.int64)
    X = np.zeros((nodes.size, MLP_IN), dtype=np.float32)
    _check(ctx._lib, ctx._lib.mpm_learn_features(ctx._ctx, mlp._p, nodes.size, lptr(nodes),
                                                  fptr(X)), "mpm_learn_features")
    return X


def learn_collect(ctx, mlp, dataset, nfine, max_rows, seed):
    _check(ctx._lib, ctx._lib.mpm_learn_collect(ctx._ctx, mlp._p, dataset._p, nfine, max_rows,
                                                 seed), "mpm_learn_collect")


def step_learned(ctx, mlp, ncoarse, nfine, guards=(0.05, 0.5)):
    """Returns the number of coarse steps that fell back to plain steps."""
    g = np.ascontiguousarray(guards, dtype=np.float64)
    fb = ctypes.c_int(0)
    _check(ctx._lib, ctx._lib.mpm_step_learned(ctx._ctx, mlp._p if mlp is not None else None,
                                                ncoarse, nfine, dptr(g), ctypes.byref(fb)),
           "mpm_step_learned")
    return fb.value


class BatchTuning(ctypes.Structure):
    """mpm_batch_tuning: work decomposition of the persistent batched step."""
    _fields_ = [("workgroups", ctypes.c_int), ("sort_chunk", ctypes.c_int),
                ("tile_chunk", ctypes.c_int), ("node_chunk", ctypes.c_int),
                ("graph_steps", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Batch:
    """A batched ensemble of independent scenes (mpm_batch_*)."""

    def __init__(self, specs_list, counts):
        self._lib = load()
        self._b = None
        n = len(specs_list)
        arr = (Specs * n)()
        for k, sp in enumerate(specs_list):
            arr[k] = sp
        self.counts = np.ascontiguousarray(counts, dtype=np.int64)
        assert self.counts.size == n
        self.nn = [(sp.ncells[0] + 1) * (sp.ncells[1] + 1) * (sp.ncells[2] + 1)
                   for sp in specs_list]
        b = _P()
        _check(self._lib, self._lib.mpm_batch_create(n, arr, lptr(self.counts), ctypes.byref(b)),
               "mpm_batch_create")
        self._b = b
        self.nscenes = n

    def close(self):
        if getattr(self, "_b", None):
            self._lib.mpm_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, name, *args):
        _check(self._lib, getattr(self._lib, "mpm_batch_" + name)(self._b, *args),
               "mpm_batch_" + name)

    def upload(self, scene, pos, rdata, idata):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        rdata = np.ascontiguousarray(rdata, dtype=np.float64)
        idata = np.ascontiguousarray(idata, dtype=np.int32)
        n = pos.shape[1]
        assert pos.shape == (3, n) and rdata.shape == (NREAL, n) and idata.shape == (NINT, n)
        self._call("upload", scene, n, dptr(pos), dptr(rdata), iptr(idata))

    def init(self):
        self._call("init")

    def step(self, n=1):
        self._call("step", n)

    def num_particles(self, scene):
        v = ctypes.c_int64(0)
        self._call("num_particles", scene, ctypes.byref(v))
        return v.value

    def download(self, scene):
        n = self.num_particles(scene)
        pos = np.zeros((3, n))
        rdata = np.zeros((NREAL, n))
        idata = np.zeros((NINT, n), dtype=np.int32)
        ids = np.zeros(n, dtype=np.int64)
        self._call("download", scene, dptr(pos), dptr(rdata), iptr(idata), lptr(ids))
        return pos, rdata, idata, ids

    def nodal(self, scene, comp):
        out = np.zeros(self.nn[scene])
        self._call("get_nodal", scene, comp, dptr(out))
        return out

    def time(self, scene):
        t = ctypes.c_double(0)
        s = ctypes.c_int64(0)
        dt = ctypes.c_double(0)
        self._call("get_time", scene, ctypes.byref(t), ctypes.byref(s), ctypes.byref(dt))
        return t.value, s.value, dt.value

    def tuning(self):
        t = BatchTuning()
        self._call("get_tuning", ctypes.byref(t))
        return t.as_dict()

    def set_tuning(self, **kw):
        t = BatchTuning()
        self._call("get_tuning", ctypes.byref(t))
        for k, v in kw.items():
            setattr(t, k, int(v))
        self._call("set_tuning", ctypes.byref(t))

    def autotune(self, trial_steps=4):
        t = BatchTuning()
        ms = ctypes.c_double(0)
        self._call("autotune", trial_steps, ctypes.byref(t), ctypes.byref(ms))
        return t.as_dict(), ms.value

    STAGES = ("keys", "scan", "scatter", "order", "p2g", "grid", "particle")

    def set_option(self, opt, value):
        self._call("set_option", opt, value)

    def profile(self, nsteps=5):
        """Per-stage workgroup-ms per step (busy, wait), tasks, other, step ms."""
        busy, wait = np.zeros(7), np.zeros(7)
        tasks = np.zeros(7, dtype=np.int64)
        other, ms = ctypes.c_double(0), ctypes.c_double(0)
        self._call("profile", nsteps, dptr(busy), dptr(wait), lptr(tasks), ctypes.byref(other),
                   ctypes.byref(ms))
        return {"stages": {k: {"busy_wg_ms": float(busy[i]), "wait_wg_ms": float(wait[i]),
                               "tasks": int(tasks[i])} for i, k in enumerate(self.STAGES)},
                "other_wg_ms": other.value, "step_ms": ms.value}

    def stats(self):
        ms = ctypes.c_double(0)
        n = ctypes.c_int64(0)
        k = ctypes.c_int(0)
        self._call("stats", ctypes.byref(ms), ctypes.byref(n), ctypes.byref(k))
        return {"step_ms": ms.value, "particles": n.value, "tasks": k.value}


def read_particle_file(path, thermal=False):
    """mpm_read_particle_file: (pos, rdata, idata, (total mass, total volume));
    thermal: mpm_read_particle_file_thermal, (pos, rdata, idata, th, totals)."""
    lib = load()
    n = ctypes.c_int64(0)
    _check(lib, lib.mpm_read_particle_file(os.fsencode(path), ctypes.byref(n), None, None, None,
                                           None), "mpm_read_particle_file")
    pos = np.zeros((3, n.value))
    rdata = np.zeros((NREAL, n.value))
    idata = np.zeros((NINT, n.value), dtype=np.int32)
    tot = np.zeros(2)
    if thermal:
        th = np.zeros((NTHERMAL, n.value))
        _check(lib, lib.mpm_read_particle_file_thermal(os.fsencode(path), ctypes.byref(n),
                                                       dptr(pos), dptr(rdata), iptr(idata),
                                                       dptr(th), dptr(tot)),
               "mpm_read_particle_file_thermal")
        return pos, rdata, idata, th, (float(tot[0]), float(tot[1]))
    _check(lib, lib.mpm_read_particle_file(os.fsencode(path), ctypes.byref(n), dptr(pos),
                                           dptr(rdata), iptr(idata), dptr(tot)),
           "mpm_read_particle_file")
    return pos, rdata, idata, (float(tot[0]), float(tot[1]))


def autogen_particles(specs, mincoords, maxcoords, vel, ppc, dens, model, E=0.0, nu=0.0, K=0.0,
                      gamma=0.0, visc=0.0, multi_per_cell=1):
    """mpm_autogen_particles: (pos, rdata, idata, (total mass, total volume))."""
    lib = load()
    mn = np.ascontiguousarray(mincoords, dtype=np.float64)
    mx = np.ascontiguousarray(maxcoords, dtype=np.float64)
    v = np.ascontiguousarray(vel, dtype=np.float64)
    pc = np.ascontiguousarray(ppc, dtype=np.int32)
    n = ctypes.c_int64(0)
    args = (ctypes.byref(specs), dptr(mn), dptr(mx), dptr(v), iptr(pc), dens, model, E, nu, K,
            gamma, visc, multi_per_cell)
    _check(lib, lib.mpm_autogen_particles(*args, ctypes.byref(n), None, None, None, None),
           "mpm_autogen_particles")
    pos = np.zeros((3, n.value))
    rdata = np.zeros((NREAL, n.value))
    idata = np.zeros((NINT, n.value), dtype=np.int32)
    tot = np.zeros(2)
    if n.value:
        _check(lib, lib.mpm_autogen_particles(*args, ctypes.byref(n), dptr(pos), dptr(rdata),
                                              iptr(idata), dptr(tot)), "mpm_autogen_particles")
    return pos, rdata, idata, (float(tot[0]), float(tot[1]))


def write_particle_file(path, pos, rdata, idata, th=None):
    """mpm_write_particle_file; with th, mpm_write_particle_file_thermal."""
    lib = load()
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    rdata = np.ascontiguousarray(rdata, dtype=np.float64)
    idata = np.ascontiguousarray(idata, dtype=np.int32)
    if th is not None:
        th = np.ascontiguousarray(th, dtype=np.float64)
        _check(lib, lib.mpm_write_particle_file_thermal(os.fsencode(path), pos.shape[1],
                                                        dptr(pos), dptr(rdata), iptr(idata),
                                                        dptr(th)),
               "mpm_write_particle_file_thermal")
        return
    _check(lib, lib.mpm_write_particle_file(os.fsencode(path), pos.shape[1], dptr(pos),
                                            dptr(rdata), iptr(idata)), "mpm_write_particle_file")


def levelset_analytic(specs, kind, params, fluid_inside=False):
    """mpm_levelset_analytic: signed distance on the nodal box, shape
    (n2+1, n1+1, n0+1) (i fastest in memory)."""
    lib = load()
    n = specs.ncells
    out = np.zeros((n[2] + 1, n[1] + 1, n[0] + 1))
    prm = np.ascontiguousarray(params, dtype=np.float64)
    _check(lib, lib.mpm_levelset_analytic(ctypes.byref(specs), kind, dptr(prm),
                                          1 if fluid_inside else 0, dptr(out)),
           "mpm_levelset_analytic")
    return out
