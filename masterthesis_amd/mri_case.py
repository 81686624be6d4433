"""Case set-up of the MRI-data brain example: the config values, time step
and initial-state file the MRIDataPBHEqnFDM program reads (--configfile,
--initfile with --initKnownDfs=1), built the way the reference's driver
examples/MRIDataPBHEqnFDM/startSimulation.py does for a synthetic case (no
MRI volume, no vessel segmentation, no PyMC fitting):

 * parse_config (startSimulation.py:34-162): sections Dimension, Geometry,
   Time, Output, Input, Brain, Tumor, Parameters; DIAMETER = 2 RADIUS;
 * derive (startSimulation.py:356-441): grid sizes, N_TIMESTEPS from the
   explicit-scheme stability estimate of brain and tumor when the config
   gives none (5 % margin), the time step, the tumor centre (middle of the
   first two axes, DEPTH below the last face) and the convergence-check
   iterates given as fractions of N_TIMESTEPS;
 * region (startSimulation.py:533-567): 1 inside the tumor sphere, else 0;
 * surface (startSimulation.py:666-713): 1 on the last-axis face inside the
   disc of radius DIAMETER * HOLE_FACTOR / 2 around the tumor centre (the
   open skull), else 0;
 * init fields (startSimulation.py:649-664, 853-903): each [Brain] value,
   replaced by the [Tumor] value inside the tumor;
 * write_init_file: classic NetCDF with the unlimited "time" dimension first
   and nNodes_0..nNodes_{d-1}; material fields and T as doubles, "surface"
   as int, each variable (time, nNodes_{d-1}, .., nNodes_0), one record.
"""
import configparser

import numpy as np

# Stefan-Boltzmann constant of the stability estimate (startSimulation.py:315;
# the update itself uses 5.670373e-8)
SIGMA_ESTIMATE = 5.670367e-8
MATERIAL_NAMES = {"RHO": "rho", "C": "c", "LAMBDA": "lambda", "RHO_BLOOD": "rho_blood", "C_BLOOD": "c_blood",
                  "OMEGA": "omega", "T_BLOOD": "T_blood", "Q": "q", "T": "T"}


def parse_config(path):
    cfg = configparser.ConfigParser()
    cfg.optionxform = str
    if not cfg.read(path):
        raise FileNotFoundError(path)
    p = {"SPACE_DIM": cfg["Dimension"].getint("SPACE_DIM", fallback=3) if cfg.has_section("Dimension") else 3}
    geo = cfg["Geometry"]
    p["COORD_NODE_FIRST"] = [float(v) for v in geo.get("COORD_NODE_FIRST").split("x")]
    p["COORD_NODE_LAST"] = [float(v) for v in geo.get("COORD_NODE_LAST").split("x")]
    p["N_NODES"] = [int(v) for v in geo.get("N_NODES").split("x")]
    t = cfg["Time"]
    p["START_TIME"] = t.getint("START_TIME", fallback=0)
    p["END_TIME"] = t.getint("END_TIME", fallback=1)
    p["N_TIMESTEPS"] = t.getint("N_TIMESTEPS", fallback=0)
    p["N_SNAPSHOTS"] = cfg["Output"].getint("N_SNAPSHOTS", fallback=1) if cfg.has_section("Output") else 1
    inp = cfg["Input"] if cfg.has_section("Input") else {}
    get = (lambda k, d: inp.get(k, d)) if isinstance(inp, dict) else (lambda k, d: inp.get(k, fallback=d))
    p["THRESHOLD"] = float(get("THRESHOLD", 0.00001))
    p["CHECK_CONV_FIRST_AT_ITER"] = float(get("CHECK_CONV_FIRST_AT_ITER", 1))
    p["CHECK_CONV_AT_EVERY_N_ITER"] = float(get("CHECK_CONV_AT_EVERY_N_ITER", 1))
    p["BRAIN"] = {k: float(v) for k, v in cfg.items("Brain")}
    p["TUMOR"] = {k: float(v) for k, v in cfg.items("Tumor")}
    par = {k: float(v) for k, v in cfg.items("Parameters")}
    if "RADIUS" in par:
        par["DIAMETER"] = 2.0 * par["RADIUS"]
    p["PARAMETERS"] = par
    return p


def delta_times(p, m):
    """(inner, border) stable time steps of one material
    (startSimulation.py:298-343)."""
    par, dim, h = p["PARAMETERS"], p["SPACE_DIM"], p["GRIDSIZE"]
    t_max = m["T"] + m["Q"] / (m["RHO_BLOOD"] * m["C_BLOOD"] * m["OMEGA"])
    tmp = 0.0
    for d in range(dim):
        tmp += (2.0 / (h[d] * h[d])) * (m["LAMBDA"] / (m["RHO"] * m["C"]))
    tmp += ((m["RHO_BLOOD"] * m["C_BLOOD"]) / (m["RHO"] * m["C"])) * m["OMEGA"]
    inner = 1.0 / tmp if tmp != 0 else float("inf")
    tmp += 2.0 * (1.0 / h[dim - 1]) * (par["H"] / (m["RHO"] * m["C"]))
    tmp += 2.0 * (1.0 / h[dim - 1]) * ((par["EPSILON"] * SIGMA_ESTIMATE) / (m["RHO"] * m["C"])) * \
        ((t_max + 273.15) ** 3)
    border = 1.0 / tmp if tmp != 0 else float("inf")
    return inner, border


def derive(p):
    """Grid sizes, N_TIMESTEPS, DELTA_TIME, TUMOR_CENTER and the convergence
    iterates (startSimulation.py:356-441); updates and returns p."""
    dim = p["SPACE_DIM"]
    p["GRIDSIZE"] = [(p["COORD_NODE_LAST"][d] - p["COORD_NODE_FIRST"][d]) / (p["N_NODES"][d] - 1)
                     for d in range(dim)]
    if p["N_TIMESTEPS"] < 1:
        dts = list(delta_times(p, p["BRAIN"])) + list(delta_times(p, p["TUMOR"]))
        p["N_TIMESTEPS"] = int(((p["END_TIME"] - p["START_TIME"]) / min(dts)) * 1.05) + 1
    p["DELTA_TIME"] = (p["END_TIME"] - p["START_TIME"]) / p["N_TIMESTEPS"]
    first, last = p["COORD_NODE_FIRST"], p["COORD_NODE_LAST"]
    p["TUMOR_CENTER"] = [(last[0] + first[0]) / 2.0, (last[1] + first[1]) / 2.0,
                         last[2] - p["PARAMETERS"]["DEPTH"]]
    for k in ("CHECK_CONV_FIRST_AT_ITER", "CHECK_CONV_AT_EVERY_N_ITER"):
        if p[k] < 1:
            p[k] = p[k] * p["N_TIMESTEPS"]
        p[k] = int(p[k])
    p["N_SNAPSHOTS"] = min(p["N_SNAPSHOTS"], p["N_TIMESTEPS"])
    return p


def _coords(p):
    nx, ny, nz = p["N_NODES"]
    h, x0 = p["GRIDSIZE"], p["COORD_NODE_FIRST"]
    x = np.arange(nx) * h[0] + x0[0]
    y = np.arange(ny) * h[1] + x0[1]
    z = np.arange(nz) * h[2] + x0[2]
    return x[None, None, :], y[None, :, None], z[:, None, None]


def region(p):
    """1 inside the tumor sphere, else 0; array (nz, ny, nx)."""
    x, y, z = _coords(p)
    c, r = p["TUMOR_CENTER"], p["PARAMETERS"]["DIAMETER"] / 2
    d = (x - c[0]) * (x - c[0])
    d = d + (y - c[1]) * (y - c[1])
    d = d + (z - c[2]) * (z - c[2])
    return (d <= r * r).astype(np.int32)


def surface(p):
    """1 on the last-axis face inside the open-skull disc, else 0."""
    x, y, _ = _coords(p)
    nx, ny, nz = p["N_NODES"]
    c = p["TUMOR_CENTER"]
    r = (p["PARAMETERS"]["DIAMETER"] * p["PARAMETERS"]["HOLE_FACTOR"]) / 2
    d = (x - c[0]) * (x - c[0]) + (y - c[1]) * (y - c[1])
    s = np.zeros((nz, ny, nx), dtype=np.int32)
    s[-1] = (d[0] <= r * r)
    return s


def init_fields(p):
    """Ordered {name: array (nz, ny, nx)}: the eight materials and T from
    [Brain], [Tumor] inside the tumor, plus "surface"."""
    reg = region(p)
    out = {"surface": surface(p)}
    for key, val in p["BRAIN"].items():
        out[MATERIAL_NAMES[key]] = np.where(reg == 1, p["TUMOR"][key], val * np.ones(reg.shape))
    return out


def write_init_file(path, fields):
    """Classic NetCDF initial-state file: "time" (unlimited) first, then
    nNodes_0..nNodes_{d-1}; one record per variable."""
    from scipy.io import netcdf_file
    shape = next(iter(fields.values())).shape
    f = netcdf_file(path, "w", version=1)
    try:
        f.createDimension("time", None)
        for d, n in enumerate(reversed(shape)):
            f.createDimension(f"nNodes_{d}", n)
        dims = ("time",) + tuple(f"nNodes_{d}" for d in reversed(range(len(shape))))
        for name, arr in fields.items():
            v = f.createVariable(name, "i" if arr.dtype.kind == "i" else "d", dims)
            v[0] = arr
    finally:
        f.close()


def command_line(p, exe, configfile, initfile, outputfile):
    """Arguments of one run (the environment the reference's driver exports,
    startSimulation.py:905-947, as options)."""
    xs = lambda v: "x".join(str(a) for a in v)  # noqa: E731
    return [exe, f"--dim={p['SPACE_DIM']}", f"--nNodes={xs(p['N_NODES'])}",
            f"--coordNodeFirst={xs(p['COORD_NODE_FIRST'])}", f"--coordNodeLast={xs(p['COORD_NODE_LAST'])}",
            f"--starttime={p['START_TIME']}", f"--endtime={p['END_TIME']}", f"--nTimesteps={p['N_TIMESTEPS']}",
            f"--nSnapshots={p['N_SNAPSHOTS']}", f"--threshold={p['THRESHOLD']}",
            f"--checkConvFirstAtIter={p['CHECK_CONV_FIRST_AT_ITER']}",
            f"--checkConvAtEveryNIter={max(1, p['CHECK_CONV_AT_EVERY_N_ITER'])}",
            f"--configfile={configfile}", f"--initfile={initfile}", "--initKnownDfs=1",
            f"--outputfile={outputfile}"]
