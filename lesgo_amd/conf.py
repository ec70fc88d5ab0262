"""The reference's input and parameter-record files.

``read_conf`` reads ``lesgo.conf`` the way ``read_input_conf`` does
(input_util.f90:53-1023): every blank (space or tab) is removed from a line
before it is looked at; a line starting with ``!`` is a comment; a block is
``NAME {`` ... ``}``; an entry is ``key=value`` with a case-insensitive key;
scalars are read as Fortran list-directed input (the first item of the
value); ``point_loc`` is a ``//``-separated list of 3-vectors and
``x/y/zplane_loc`` comma-separated lists.  Unknown blocks and keys are
reported and skipped, as the reference does; a malformed entry raises
``ConfError`` (the reference stops).  The values each block derives when it
closes follow the reference: the vertical sizes and grid spacings (DOMAIN,
with one process owning all levels), ``dt_dim`` (TIME) and the evaluated
mean pressure force (FLOW_COND).

``LesgoConf.to_config`` maps the result onto this package's ``Config`` and
rejects what the time step here does not implement (sgs_model != 5, ifilter
!= 1, lbc_mom != 2, ubc_mom != 0, inflow_type != 0, Coriolis, sponge, random
forcing, DNS, rotating disks).  ``place_turbines`` is turbines.f90:156-175,
722-869 (orientations 1-5 and ``input_turbines/param.dat``);
``read_control_files`` is turbines.f90:872-931.  ``write_param_out`` writes
``output/lesgo_param.out`` with the records of param_output.f90.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

PI = 3.1415926535897932384626433


class ConfError(ValueError):
    pass


# --------------------------------------------------------------------------
# Fortran list-directed scalars
# --------------------------------------------------------------------------
def _first_item(s: str) -> str:
    for sep in (",", "/"):
        i = s.find(sep)
        if i >= 0:
            s = s[:i]
    return s


def _int(s: str) -> int:
    t = _first_item(s)
    try:
        return int(t)
    except ValueError:
        raise ConfError(f"integer expected, got {s!r}") from None


def _real(s: str) -> float:
    t = _first_item(s).replace("d", "e").replace("D", "e")
    try:
        return float(t)
    except ValueError:
        raise ConfError(f"real expected, got {s!r}") from None


def _logical(s: str) -> bool:
    t = _first_item(s).lstrip(".").upper()
    if t[:1] == "T":
        return True
    if t[:1] == "F":
        return False
    raise ConfError(f"logical expected, got {s!r}")


def _string(s: str) -> str:
    t = s.strip()
    if len(t) >= 2 and t[0] == t[-1] and t[0] in "'\"":
        return t[1:-1]
    return _first_item(t)


def _vector(s: str) -> list:
    return [_real(x) for x in s.split(",") if x != ""]


def _points(s: str) -> list:
    out = []
    for part in s.split("//"):
        v = _vector(part)
        if len(v) != 3:
            raise ConfError("vector not specified correctly")
        out.append(tuple(v))
    return out


# key -> (attribute, reader) per block (input_util.f90:133-815; LEVEL_SET and
# SCALARS belong to builds with level sets / scalars and are skipped here)
_BLOCKS = {
    "DOMAIN": {
        "NPROC": ("nproc", _int), "NX": ("nx", _int), "NY": ("ny", _int),
        "NZ": ("nz_tot", _int), "Z_I": ("z_i", _real), "LX": ("L_x", _real),
        "LY": ("L_y", _real), "LZ": ("L_z", _real),
        "UNIFORM_SPACING": ("uniform_spacing", _logical),
    },
    "MODEL": {
        "SGS_MODEL": ("sgs_model", _int), "WALL_DAMP_EXP": ("wall_damp_exp", _int),
        "CS_COUNT": ("cs_count", _int), "DYN_INIT": ("dyn_init", _int), "CO": ("Co", _real),
        "IFILTER": ("ifilter", _int), "U_STAR": ("u_star", _real), "VONK": ("vonk", _real),
        "NU_MOLEC": ("nu_molec", _real), "MOLEC": ("molec", _logical),
        "SGS": ("sgs", _logical),
    },
    "CORIOLIS": {
        "CORIOLIS_FORCING": ("coriolis_forcing", _int), "FC": ("fc", _real),
        "G": ("G", _real), "ALPHA": ("alpha", _real), "PID_TIME": ("pid_time", _int),
        "PHI_SET": ("phi_set", _real), "HEIGHT_SET": ("height_set", _real),
        "KP": ("Kp", _real), "KI": ("Ki", _real), "KD": ("Kd", _real),
        "REPEAT_INTERVAL": ("repeat_interval", _real),
    },
    "TIME": {
        "NSTEPS": ("nsteps", _int), "RUNTIME": ("runtime", _int),
        "USE_CFL_DT": ("use_cfl_dt", _logical), "CFL": ("cfl", _real), "DT": ("dt", _real),
        "CUMULATIVE_TIME": ("cumulative_time", _logical),
    },
    "FLOW_COND": {
        "INITU": ("initu", _logical), "INILAG": ("inilag", _logical),
        "LBC_MOM": ("lbc_mom", _int), "UBC_MOM": ("ubc_mom", _int), "UBOT": ("ubot", _real),
        "UTOP": ("utop", _real), "USE_SPONGE": ("use_sponge", _logical),
        "SPONGE_FREQUENCY": ("sponge_frequency", _real),
        "SPONGE_HEIGHT": ("sponge_height", _real), "ZO": ("zo", _real),
        "INFLOW_TYPE": ("inflow_type", _int),
        "FRINGE_REGION_END": ("fringe_region_end", _real),
        "FRINGE_REGION_LEN": ("fringe_region_len", _real),
        "INFLOW_VELOCITY": ("inflow_velocity", _real),
        "SAMPLING_REGION_END": ("sampling_region_end", _real), "SHIFT_N": ("shift_n", _int),
        "USE_MEAN_P_FORCE": ("use_mean_p_force", _logical),
        "EVAL_MEAN_P_FORCE": ("eval_mean_p_force", _logical),
        "MEAN_P_FORCE_X": ("mean_p_force_x", _real),
        "MEAN_P_FORCE_Y": ("mean_p_force_y", _real),
        "USE_RANDOM_FORCE": ("use_random_force", _logical),
        "STOP_RANDOM_FORCE": ("stop_random_force", _int),
        "RMS_RANDOM_FORCE": ("rms_random_force", _real),
    },
    "OUTPUT": {
        "WBASE": ("wbase", _int), "NENERGY": ("nenergy", _int),
        "LAG_CFL_COUNT": ("lag_cfl_count", _int),
        "CHECKPOINT_DATA": ("checkpoint_data", _logical),
        "CHECKPOINT_NSKIP": ("checkpoint_nskip", _int),
        "TAVG_CALC": ("tavg_calc", _logical), "TAVG_NSTART": ("tavg_nstart", _int),
        "TAVG_NEND": ("tavg_nend", _int), "TAVG_NSKIP": ("tavg_nskip", _int),
        "POINT_CALC": ("point_calc", _logical), "POINT_NSTART": ("point_nstart", _int),
        "POINT_NEND": ("point_nend", _int), "POINT_NSKIP": ("point_nskip", _int),
        "POINT_LOC": ("point_loc", _points),
        "DOMAIN_CALC": ("domain_calc", _logical), "DOMAIN_NSTART": ("domain_nstart", _int),
        "DOMAIN_NEND": ("domain_nend", _int), "DOMAIN_NSKIP": ("domain_nskip", _int),
        "XPLANE_CALC": ("xplane_calc", _logical), "XPLANE_NSTART": ("xplane_nstart", _int),
        "XPLANE_NEND": ("xplane_nend", _int), "XPLANE_NSKIP": ("xplane_nskip", _int),
        "XPLANE_LOC": ("xplane_loc", _vector),
        "YPLANE_CALC": ("yplane_calc", _logical), "YPLANE_NSTART": ("yplane_nstart", _int),
        "YPLANE_NEND": ("yplane_nend", _int), "YPLANE_NSKIP": ("yplane_nskip", _int),
        "YPLANE_LOC": ("yplane_loc", _vector),
        "ZPLANE_CALC": ("zplane_calc", _logical), "ZPLANE_NSTART": ("zplane_nstart", _int),
        "ZPLANE_NEND": ("zplane_nend", _int), "ZPLANE_NSKIP": ("zplane_nskip", _int),
        "ZPLANE_LOC": ("zplane_loc", _vector),
    },
    "TURBINES": {
        "NUM_X": ("num_x", _int), "NUM_Y": ("num_y", _int), "DIA_ALL": ("dia_all", _real),
        "HEIGHT_ALL": ("height_all", _real), "THK_ALL": ("thk_all", _real),
        "ORIENTATION": ("orientation", _int), "STAG_PERC": ("stag_perc", _real),
        "THETA1_ALL": ("theta1_all", _real), "THETA2_ALL": ("theta2_all", _real),
        "CT_PRIME": ("Ct_prime", _real), "READ_PARAM": ("read_param", _logical),
        "DYN_THETA1": ("dyn_theta1", _logical), "DYN_THETA2": ("dyn_theta2", _logical),
        "DYN_CT_PRIME": ("dyn_Ct_prime", _logical),
        "USE_ROTATION": ("use_rotation", _logical),
        "TIP_SPEED_RATIO": ("tip_speed_ratio", _real), "T_AVG_DIM": ("T_avg_dim", _real),
        "ALPHA1": ("alpha1", _real), "ALPHA2": ("alpha2", _real),
        "FILTER_CUTOFF": ("filter_cutoff", _real),
        "ADM_CORRECTION": ("adm_correction", _logical), "TBASE": ("tbase", _int),
    },
}


@dataclass
class LesgoConf:
    """param.f90 (and the turbines module) after reading lesgo.conf; defaults
    are param.f90:96-266.  ``turbines`` is True when a TURBINES block was read
    (the PPTURBINES build reads it; its values have no defaults)."""
    nproc: int = 1
    nx: int = 64
    ny: int = 64
    nz_tot: int = 64
    z_i: float = 1000.0
    L_x: float = 2.0 * PI
    L_y: float = 2.0 * PI
    L_z: float = 1.0
    uniform_spacing: bool = False
    sgs_model: int = 5
    wall_damp_exp: int = 2
    cs_count: int = 5
    dyn_init: int = 100
    Co: float = 0.16
    ifilter: int = 1
    u_star: float = 0.45
    vonk: float = 0.4
    nu_molec: float = 1.14e-5
    molec: bool = False
    sgs: bool = True
    coriolis_forcing: int = 0
    fc: float = 1.0e-4
    G: float = 8.0
    alpha: float = 0.0
    pid_time: int = 1000000
    phi_set: float = 0.0
    height_set: float = 100.0
    Kp: float = 0.0
    Ki: float = 0.0
    Kd: float = 0.0
    repeat_interval: float = 86400.0
    nsteps: int = 50000
    runtime: int = -1
    use_cfl_dt: bool = False
    cfl: float = 0.05
    dt: float = 2.0e-4
    cumulative_time: bool = True
    initu: bool = False
    inilag: bool = True
    lbc_mom: int = 1
    ubc_mom: int = 0
    ubot: float = 0.0
    utop: float = 0.0
    use_sponge: bool = False
    sponge_frequency: float = 3.9
    sponge_height: float = 1.5
    zo: float = 1.0e-4
    inflow_type: int = 0
    fringe_region_end: float = 1.0
    fringe_region_len: float = 0.125
    inflow_velocity: float = 1.0
    sampling_region_end: float = 0.625
    shift_n: int = 1
    use_mean_p_force: bool = True
    eval_mean_p_force: bool = False
    mean_p_force_x: float = 1.0
    mean_p_force_y: float = 0.0
    use_random_force: bool = False
    stop_random_force: int = 20000
    rms_random_force: float = 0.4
    wbase: int = 100
    nenergy: int = 100
    lag_cfl_count: int = 1000
    checkpoint_data: bool = False
    checkpoint_nskip: int = 10000
    tavg_calc: bool = False
    tavg_nstart: int = 1
    tavg_nend: int = 50000
    tavg_nskip: int = 100
    point_calc: bool = False
    point_nstart: int = 1
    point_nend: int = 50000
    point_nskip: int = 10
    point_loc: list = field(default_factory=lambda: [(0.0, 0.0, 0.0)])
    domain_calc: bool = False
    domain_nstart: int = 10000
    domain_nend: int = 50000
    domain_nskip: int = 10000
    xplane_calc: bool = False
    xplane_nstart: int = 10000
    xplane_nend: int = 50000
    xplane_nskip: int = 10000
    xplane_loc: list = field(default_factory=lambda: [0.0])
    yplane_calc: bool = False
    yplane_nstart: int = 10000
    yplane_nend: int = 50000
    yplane_nskip: int = 10000
    yplane_loc: list = field(default_factory=lambda: [0.0])
    zplane_calc: bool = False
    zplane_nstart: int = 10000
    zplane_nend: int = 50000
    zplane_nskip: int = 10000
    zplane_loc: list = field(default_factory=lambda: [0.0])
    # turbines module
    turbines: bool = False
    num_x: int = 0
    num_y: int = 0
    dia_all: float = 0.0
    height_all: float = 0.0
    thk_all: float = 0.0
    orientation: int = 1
    stag_perc: float = 0.0
    theta1_all: float = 0.0
    theta2_all: float = 0.0
    Ct_prime: float = 0.0
    read_param: bool = False
    dyn_theta1: bool = False
    dyn_theta2: bool = False
    dyn_Ct_prime: bool = False
    use_rotation: bool = False
    tip_speed_ratio: float = 7.0
    T_avg_dim: float = -1.0
    alpha1: float = 0.0
    alpha2: float = 0.0
    filter_cutoff: float = 0.0
    adm_correction: bool = False
    tbase: int = 20
    # derived (input_util.f90:197-235, 406-409)
    nz: int = 0
    dx: float = 0.0
    dy: float = 0.0
    dz: float = 0.0
    dt_dim: float = 0.0
    path: str = "./"
    messages: list = field(default_factory=list)

    # -- derived values when a block closes --------------------------------
    def _close_domain(self):
        # one process owns every level (the reference's non-MPI branch)
        if self.nproc != 1:
            self.messages.append(f"Reseting nproc to: 1 (one GPU holds all {self.nz_tot} levels)")
            self.nproc = 1
        self.nz = self.nz_tot // self.nproc + 1
        old = self.nz_tot
        self.nz_tot = (self.nz - 1) * self.nproc + 1
        if old != self.nz_tot:
            self.messages.append(f"Reseting Nz (total) to: {self.nz_tot}")
        self.dx = self.L_x / self.nx
        if self.uniform_spacing:
            self.L_y = self.ny * self.dx
            self.L_z = (self.nz_tot - 1) * self.dx
        self.dy = self.L_y / self.ny
        self.dz = self.L_z / (self.nz_tot - 1)

    def _close_time(self):
        if not self.use_cfl_dt:
            self.dt_dim = self.dt * self.z_i / self.u_star

    def _close_flow_cond(self):
        if self.use_mean_p_force and self.eval_mean_p_force:
            # the y component uses the already updated x component, as the
            # reference does (input_util.f90:525-527)
            self.mean_p_force_x = (self.mean_p_force_x
                                   / math.sqrt(self.mean_p_force_x ** 2
                                               + self.mean_p_force_y ** 2) / self.L_z)
            self.mean_p_force_y = (self.mean_p_force_y
                                   / math.sqrt(self.mean_p_force_x ** 2
                                               + self.mean_p_force_y ** 2) / self.L_z)

    @property
    def Nz(self) -> int:
        """uv levels of the one-process grid (the ABI's Nz)."""
        return self.nz - 1

    @property
    def nu(self) -> float:
        """sgs_param.f90:187-192."""
        return self.nu_molec / (self.u_star * self.z_i) if self.molec else 0.0

    def check_supported(self):
        """Settings outside the implemented time step raise ConfError."""
        bad = []
        if self.sgs_model != 5:
            bad.append(f"sgs_model = {self.sgs_model} (only 5, Lagrangian scale-dependent)")
        if self.ifilter != 1:
            bad.append(f"ifilter = {self.ifilter} (only 1, spectral cut-off)")
        if self.lbc_mom != 2:
            bad.append(f"lbc_mom = {self.lbc_mom} (only 2, equilibrium wall model)")
        if self.ubc_mom != 0:
            bad.append(f"ubc_mom = {self.ubc_mom} (only 0, stress-free lid)")
        if self.inflow_type != 0:
            bad.append(f"inflow_type = {self.inflow_type} (only 0, periodic)")
        if self.coriolis_forcing != 0:
            bad.append("coriolis_forcing != 0")
        if self.use_sponge:
            bad.append("use_sponge = .true.")
        if self.use_random_force:
            bad.append("use_random_force = .true.")
        if not self.sgs:
            bad.append("sgs = .false. (DNS)")
        if self.turbines and self.use_rotation:
            bad.append("use_rotation = .true.")
        if bad:
            raise ConfError("not supported by this time step: " + "; ".join(bad))

    def to_config(self, precision: str = "fp64", initu: bool | None = None):
        """The package's Config for this run (initu: from vel.out presence
        when None, initial.f90:96-105)."""
        from .case import Config
        self.check_supported()
        if initu is None:
            initu = os.path.exists(os.path.join(self.path, "vel.out"))
        return Config(self.nx, self.ny, self.Nz, lx=self.L_x, ly=self.L_y, lz=self.L_z,
                      dt=self.dt, zo=self.zo, vonk=self.vonk, cs_count=self.cs_count,
                      use_mean_p_force=self.use_mean_p_force,
                      mean_p_force_x=self.mean_p_force_x, mean_p_force_y=self.mean_p_force_y,
                      nu=self.nu, precision=precision, initu=initu, dyn_init=self.dyn_init,
                      use_cfl_dt=self.use_cfl_dt, cfl=self.cfl)


def _clean(line: str) -> str:
    return line.replace(" ", "").replace("\t", "").rstrip("\r\n")


def read_conf(path: str = "./", text: str | None = None) -> LesgoConf:
    """Read ``path/lesgo.conf`` (or ``text``) into a LesgoConf."""
    if text is None:
        fname = os.path.join(path, "lesgo.conf")
        if not os.path.exists(fname):
            raise ConfError(f"file {fname} does not exist")
        with open(fname) as f:
            text = f.read()
    c = LesgoConf(path=path)
    lines = [(n + 1, _clean(raw)) for n, raw in enumerate(text.splitlines())]
    lines = [(n, b) for n, b in lines if b and not b.startswith("!")]
    i = 0
    while i < len(lines):
        n, b = lines[i]
        be = b.find("{")
        if be < 0:
            raise ConfError(f"block entry not found on line {n}")
        name = b[:be].upper()
        keys = _BLOCKS.get(name)
        i += 1
        if keys is None:
            c.messages.append(f"Found unused input block: {b[:be]}")
            while i < len(lines) and "}" not in lines[i - 1][1]:
                i += 1
            continue
        if name == "TURBINES":
            c.turbines = True
        closed = False
        while i < len(lines):
            n, b = lines[i]
            i += 1
            bx = b.find("}")
            if bx == 0:
                closed = True
                break
            if bx > 0:
                eq = b.find("=")
                raise ConfError(f"{name} data block not formatted correctly: "
                                f"{b[:eq] if eq >= 0 else b}")
            eq = b.find("=")
            if eq < 0:
                raise ConfError(f"Bad read in block at line {n}: {b}")
            if eq == len(b) - 1:
                raise ConfError(f"nothing after equals sign in line {n}")
            key = b[:eq].upper()
            ent = keys.get(key)
            if ent is None:
                c.messages.append(f"Found unused data value in {name} block: {b[:eq]}")
                continue
            attr, reader = ent
            try:
                setattr(c, attr, reader(b[eq + 1:]))
            except ConfError as e:
                raise ConfError(f"line {n} ({key}): {e}") from None
        if not closed:
            raise ConfError(f"Bad read in block {name}")
        if name == "DOMAIN":
            c._close_domain()
        elif name == "TIME":
            c._close_time()
        elif name == "FLOW_COND":
            c._close_flow_cond()
    if c.nz == 0:
        c._close_domain()
    return c


# --------------------------------------------------------------------------
# Turbine placement (turbines.f90:156-175, 722-869) and control schedules
# --------------------------------------------------------------------------
def place_turbines(c: LesgoConf):
    """Rows (xloc, yloc, height, dia, thk, theta1, theta2, Ct') in units of
    z_i, turbine k = (i - 1) num_y + j for row i of num_x, column j of
    num_y."""
    nloc = c.num_x * c.num_y
    if c.read_param:
        fname = os.path.join(c.path, "input_turbines", "param.dat")
        if not os.path.exists(fname):
            raise ConfError(f"file {fname} does not exist")
        rows = [ln for ln in open(fname).read().splitlines() if ln.strip()]
        if len(rows) < nloc:
            raise ConfError(f"{fname} must have num_x*num_y lines")
        if len(rows) > nloc:
            c.messages.append(f"{fname} has more than num_x*num_y lines. "
                              "Only reading first num_x*num_y lines")
        out = []
        for ln in rows[:nloc]:
            v = [_real(x) for x in ln.replace(",", " ").split()]
            if len(v) < 8:
                raise ConfError(f"{fname}: 8 values per line expected")
            x, y, h, d, t = (q / c.z_i for q in v[:5])
            out.append((x, y, h, d, t, v[5], v[6], v[7]))
        return out
    h_all, d_all, t_all = c.height_all / c.z_i, c.dia_all / c.z_i, c.thk_all / c.z_i
    sx = c.L_x / (c.num_x * d_all)
    sy = c.L_y / (c.num_y * d_all)
    sxx, syy = sx * d_all, sy * d_all
    t = []
    for i in range(1, c.num_x + 1):
        for j in range(1, c.num_y + 1):
            t.append([sxx * (2 * i - 1) / 2, syy * (2 * j - 1) / 2, h_all, d_all, t_all,
                      c.theta1_all, c.theta2_all, c.Ct_prime])
    o = c.orientation
    if o == 1:
        pass
    elif o == 2:
        shift_base = syy * c.stag_perc / 100.0
        for i in range(2, c.num_x + 1):
            for k in range(c.num_y * (i - 1), c.num_y * i):
                t[k][1] = math.fmod(t[k][1] + (i - 1) * shift_base, c.L_y)
    elif o == 3:
        for i in range(1, c.num_x + 1):
            fac = (1.0 + c.stag_perc / 100.0) if i % 2 == 0 else (1.0 - c.stag_perc / 100.0)
            for k in range(c.num_y * (i - 1), c.num_y * i):
                t[k][2] = h_all * fac
    elif o == 4:
        k = 0
        for i in range(1, c.num_x + 1):
            for j in range(1, c.num_y + 1):
                const = 2.0 * math.fmod(float(i + j), 2.0) - 1.0
                t[k][2] = h_all * (1.0 + const * c.stag_perc / 100.0)
                k += 1
    elif o == 5:
        dummy = c.stag_perc * (t[1][1] - t[0][1]) if nloc > 1 else 0.0
        k = 0
        for i in range(1, c.num_x + 1):
            for j in range(1, c.num_y + 1):
                t[k][1] = math.fmod(t[k][1] + dummy * (i - 1), c.L_y)
                k += 1
    else:
        raise ConfError("invalid orientation")
    return [tuple(r) for r in t]


def read_control_files(c: LesgoConf):
    """Dynamic theta1 / theta2 / Ct' schedules (turbines.f90:872-931): for
    each enabled one, (times, values[turbine][time]) from
    input_turbines/<name>.dat, one line per time point."""
    out = {}
    nloc = c.num_x * c.num_y
    for flag, name in ((c.dyn_theta1, "theta1"), (c.dyn_theta2, "theta2"),
                       (c.dyn_Ct_prime, "Ct_prime")):
        if not flag:
            continue
        fname = os.path.join(c.path, "input_turbines", f"{name}.dat")
        if not os.path.exists(fname):
            raise ConfError(f"file {fname} does not exist")
        times, vals = [], []
        for ln in open(fname).read().splitlines():
            if not ln.strip():
                continue
            v = [_real(x) for x in ln.replace(",", " ").split()]
            if len(v) < nloc + 1:
                raise ConfError(f"{fname}: time and {nloc} values per line expected")
            times.append(v[0])
            vals.append(v[1:nloc + 1])
        out[name] = (times, [[row[s] for row in vals] for s in range(nloc)])
    return out


def linear_interp(x, v, xq):
    """functions.f90:712-735 with binary_search 811-860: v(1) below x(1),
    v(N) above x(N), linear in between."""
    n = len(x)
    if xq < x[0]:
        return v[0]
    if xq > x[-1]:
        return v[-1]
    lo, hi = 0, n - 1
    while hi - lo > 1:
        mid = (lo + hi) // 2
        if x[mid] > xq:
            hi = mid
        elif x[mid] < xq:
            lo = mid
        else:
            lo = mid
            break
    if lo == n - 1:
        return v[-1]
    return v[lo] + (xq - x[lo]) * (v[lo + 1] - v[lo]) / (x[lo + 1] - x[lo])


# --------------------------------------------------------------------------
# output/lesgo_param.out (param_output.f90)
# --------------------------------------------------------------------------
def _e15(x: float) -> str:
    """Fortran E15.7: sign, 0.ddddddd, E, signed two-digit exponent,
    right-justified in 15 columns."""
    if x == 0.0:
        return "0.0000000E+00".rjust(15)
    m, e = f"{abs(x):.6e}".split("e")
    exp = int(e) + 1
    s = f"0.{m.replace('.', '')}E{'+' if exp >= 0 else '-'}{abs(exp):02d}"
    return ("-" + s if x < 0 else s).rjust(15)


def _i7(n: int) -> str:
    return f"{n:7d}"


def _l(b: bool) -> str:
    return " T" if b else " F"


def write_param_out(c: LesgoConf, fname: str | None = None):
    """output/lesgo_param.out with the records of param_output.f90 (PPCFL_DT
    off: dt is written; endianness NATIVE)."""
    if fname is None:
        os.makedirs(os.path.join(c.path, "output"), exist_ok=True)
        fname = os.path.join(c.path, "output", "lesgo_param.out")
    star = "*" * 70
    dash = "-" * 51
    L = []
    L += [star, "PARAM", star, "", dash, "COMPUTATIONAL DOMAIN PARAMETERS", dash]
    L.append("nproc : " + _i7(c.nproc))
    L.append("nx, ny, nz, nz_tot : " + "".join(_i7(q) for q in (c.nx, c.ny, c.nz, c.nz_tot)))
    L.append("z_i : " + _e15(c.z_i))
    L.append("L_x, L_y, L_z : " + "".join(_e15(q) for q in (c.L_x, c.L_y, c.L_z)))
    L.append("dx, dy, dz : " + "".join(_e15(q) for q in (c.dx, c.dy, c.dz)))
    L += ["", dash, "MODEL PARAMETERS", dash]
    L.append("sgs_model, wall_damp_exp : " + _i7(c.sgs_model) + _i7(c.wall_damp_exp))
    L.append("Co : " + _e15(c.Co))
    L.append("cs_count : " + _i7(c.cs_count))
    L.append("ifilter : " + _i7(c.ifilter))
    L.append("u_star : " + _e15(c.u_star))
    L.append("vonk : " + _e15(c.vonk))
    L.append("nu_molec : " + _e15(c.nu_molec))
    L.append("molec, sgs : " + _l(c.molec) + _l(c.sgs))
    L += ["", dash, "TIMESTEP PARAMETERS", dash]
    L.append("nsteps : " + _i7(c.nsteps))
    L.append("runtime : " + _i7(c.runtime))
    L.append("dt : " + _e15(c.dt))
    L.append("cumulative_time : " + _l(c.cumulative_time))
    L += ["", dash, "BOUNDARY/INITIAL CONDITION PARAMETERS", dash]
    L.append("initu : " + _l(c.initu))
    L.append("inilag : " + _l(c.inilag))
    L.append("lbc_mom : " + _i7(c.lbc_mom))
    L.append("ubc_mom : " + _i7(c.ubc_mom))
    L.append("ubot : " + _e15(c.ubot))
    L.append("utop : " + _e15(c.utop))
    L.append("zo : " + _e15(c.zo))
    L.append("inflow_type : " + _i7(c.inflow_type))
    L.append("fringe_region_end : " + _e15(c.fringe_region_end))
    L.append("fringe_region_len : " + _e15(c.fringe_region_len))
    L.append("inflow_velocity : " + _e15(c.inflow_velocity))
    L.append("use_mean_p_force : " + _l(c.use_mean_p_force))
    L.append("mean_p_force_x : " + _e15(c.mean_p_force_x))
    L.append("mean_p_force_y : " + _e15(c.mean_p_force_y))
    L.append("use_random_force : " + _l(c.use_random_force))
    L.append("stop_random_force : " + _i7(c.stop_random_force))
    L.append("rms_random_force : " + _e15(c.rms_random_force))
    L += ["", dash, "DATA OUTPUT PARAMETERS", dash]
    L.append("write_endian : NATIVE")
    L.append("read_endian : NATIVE")
    L.append("wbase : " + _i7(c.wbase))
    L.append("nenergy : " + _i7(c.nenergy))
    L.append("lag_cfl_count : " + _i7(c.lag_cfl_count))
    L.append("checkpoint_data : " + _l(c.checkpoint_data))
    L.append("checkpoint_nskip : " + _i7(c.checkpoint_nskip))
    L.append("tavg_calc : " + _l(c.tavg_calc))
    L.append("tavg_nstart, tavg_nend, tavg_nskip : "
             + "".join(_i7(q) for q in (c.tavg_nstart, c.tavg_nend, c.tavg_nskip)))
    L.append("point_calc : " + _l(c.point_calc))
    L.append("point_nstart, point_nend, point_nskip : "
             + "".join(_i7(q) for q in (c.point_nstart, c.point_nend, c.point_nskip)))
    L.append("point_nloc : " + _i7(len(c.point_loc)))
    for n, xyz in enumerate(c.point_loc, 1):
        L.append("n, point_loc(n)%xyz : " + _i7(n) + "".join(_e15(q) for q in xyz))
    L.append("domain_calc : " + _l(c.domain_calc))
    L.append("domain_nstart, domain_nend, domain_nskip : "
             + "".join(_i7(q) for q in (c.domain_nstart, c.domain_nend, c.domain_nskip)))
    for p in ("x", "y", "z"):
        L.append(f"{p}plane_calc : " + _l(getattr(c, f"{p}plane_calc")))
        L.append(f"{p}plane_nstart, {p}plane_nend, {p}plane_nskip : "
                 + "".join(_i7(getattr(c, f"{p}plane_{q}")) for q in ("nstart", "nend", "nskip")))
        locs = getattr(c, f"{p}plane_loc")
        L.append(f"{p}plane_nloc : " + _i7(len(locs)))
        for n, v in enumerate(locs, 1):
            L.append(f"n, {p}plane_loc(n) : " + _i7(n) + _e15(v))
    with open(fname, "w") as f:
        f.write("\n".join(L) + "\n")
    return fname
