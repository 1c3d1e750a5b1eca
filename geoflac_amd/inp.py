"""Reader for GeoFLAC's positional `.inp` parameter files.

Follows src/read_params.f90:3-373: values are read in a fixed order; before
most reads, lines whose first character is ';' are skipped
(AdvanceToNextInputLine, src/read_params.f90:351-373); each read is
list-directed (comma or blank separated, continuing onto following lines if
a line holds too few values, the rest of the last line discarded).  Unit
conversions are the reference's: times in kyr -> s, bottom heat flux
mW/m2 -> W/m2, weaken_ratio_viscous -> its logarithm.
"""
import math

from . import abi

SEC_YEAR = 3.1558e7  # src/params.f90:15


class InpError(ValueError):
    pass


class _Reader:
    def __init__(self, text):
        self.lines = text.splitlines()
        self.pos = 0

    def advance(self):
        """Skip comment lines (first character ';'), src/read_params.f90:351-373."""
        while True:
            if self.pos >= len(self.lines):
                raise InpError("EOF reached while skipping comments (reference: STOP 11)")
            ln = self.lines[self.pos]
            if ln[:1] != ";":
                return
            self.pos += 1

    def read(self, n, conv=float):
        """List-directed read of n values; returns a list."""
        out = []
        while len(out) < n:
            if self.pos >= len(self.lines):
                raise InpError("EOF in the middle of a read (reference: STOP 11)")
            toks = self.lines[self.pos].replace(",", " ").split()
            self.pos += 1
            for t in toks:
                if len(out) == n:
                    break
                out.append(t)
        vals = []
        for t, c in zip(out, conv if isinstance(conv, (list, tuple)) else [conv] * n):
            try:
                if c is int:
                    vals.append(int(t))
                elif c is float:
                    vals.append(float(t.replace("d", "e").replace("D", "e")))
                else:
                    vals.append(t)
            except ValueError:
                raise InpError("cannot read %r as %s at line %d" % (t, c.__name__, self.pos))
        return vals

    def skip_record(self):
        self.pos += 1


def parse_inp(text):
    """Parse `.inp` text into a dict of the reference's parameter names."""
    r = _Reader(text)
    p = {}
    r.advance(); nx, nz = r.read(2, int)
    p["nelx"], p["nelz"] = nx, nz
    p["nx"], p["nz"] = nx + 1, nz + 1
    r.advance(); p["x0"], p["z0"] = r.read(2)
    r.advance(); p["rxbo"], p["rzbo"] = r.read(2)
    r.advance(); p["ircoord"], p["coordfile"] = r.read(2, [int, str])
    if p["ircoord"] > 0:
        r.advance(); p["nzonx"] = r.read(1, int)[0]
        r.advance(); p["nzony"] = r.read(1, int)[0]
        p["nelz_x"], p["sizez_x"], p["nelz_y"], p["sizez_y"] = [], [], [], []
    else:
        r.advance(); nzonx = r.read(1, int)[0]
        if nzonx == 0:
            p["nzonx"], p["nelz_x"], p["sizez_x"] = 1, [nx], [1.0]
        else:
            p["nzonx"], p["nelz_x"], p["sizez_x"] = nzonx, [], []
            for _ in range(nzonx):
                r.advance(); a, b = r.read(2, [int, float])
                p["nelz_x"].append(a); p["sizez_x"].append(b)
        r.advance(); nzony = r.read(1, int)[0]
        if nzony == 0:
            p["nzony"], p["nelz_y"], p["sizez_y"] = 1, [nz], [1.0]
        else:
            p["nzony"], p["nelz_y"], p["sizez_y"] = nzony, [], []
            for _ in range(nzony):
                r.advance(); a, b = r.read(2, [int, float])
                p["nelz_y"].append(a); p["sizez_y"].append(b)
    # mechanical conditions
    r.advance(); p["nystressbc"], p["nydrsides"] = r.read(2, int)
    r.advance(); p["nofbc"] = r.read(1, int)[0]
    r.advance()
    p["bcs"] = []
    for _ in range(p["nofbc"]):
        v = r.read(13, [int, int, int, int] + [float] * 9)
        p["bcs"].append(dict(zip(["nofside", "nbc1", "nbc2", "nbc", "bca", "bcb", "bcc", "bcd",
                                  "bce", "bcf", "bcg", "bch", "bci"], v)))
    r.advance()
    r.advance()
    (p["nyhydro"], p["pisos"], p["iphsub"], p["drosub"],
     p["damp_vis"]) = r.read(5, [int, float, int, float, float])
    r.advance(); p["g"] = r.read(1)[0]
    # thermal conditions
    r.advance(); p["i_prestress"] = r.read(1, int)[0]
    r.advance(); p["extra_pres"] = r.read(1)[0]
    r.advance(); p["itherm"] = r.read(1, int)[0]
    r.advance(); p["istress_therm"] = r.read(1, int)[0]
    r.advance(); p["ishearh"] = r.read(1, int)[0]
    r.advance(); p["t_top"] = r.read(1)[0]
    r.advance(); p["t_bot"] = r.read(1)[0]
    r.advance(); p["hs"], p["hr"] = r.read(2)
    r.advance(); p["itemp_bc"], p["bot_bc"] = r.read(2, [int, float])
    if p["itemp_bc"] == 2:
        p["bot_bc"] = p["bot_bc"] / 1000
    r.advance(); p["irtemp"] = r.read(1, int)[0]
    r.advance()
    if p["irtemp"] > 0:
        p["tempfile"] = r.read(1, str)[0]
    else:
        r.skip_record()
    r.advance(); p["nzone_age"] = r.read(1, int)[0]
    r.advance()
    zones = []
    for _ in range(p["nzone_age"]):
        ict, a1, t1, t2, x1, x2 = r.read(6, [int, float, float, float, int, int])
        r.advance()
        nl = r.read(1, int)[0]
        # nph_layer and hc share one record: re-read it with the right count
        r.pos -= 1
        vals = r.read(nl, [int] + [float] * (nl - 1))
        hc = vals[1:]
        iph = r.read(nl, int)
        zones.append(dict(ictherm=ict, age_1=a1, tp1=t1, tp2=t2, ixtb1=x1, ixtb2=x2,
                          nph_layer=nl, hc=hc, iph_col=iph, iph_col_trans=0))
    for k in range(len(zones) - 1):
        if zones[k]["ixtb2"] == -1:
            zones[k]["iph_col_trans"] = 1
            zones[k]["ixtb2"] = zones[k + 1]["ixtb2"]
            zones[k + 1]["ixtb1"] = zones[k]["ixtb1"]
    p["zones"] = zones
    # rheology
    r.advance(); p["nphase"] = r.read(1, int)[0]
    names = ["irheol", "visc", "den", "alfa", "beta", "pln", "acoef", "eactiv", "vactiv", "rl",
             "rm", "plstrain1", "plstrain2", "fric1", "fric2", "cohesion1", "cohesion2",
             "dilat1", "dilat2", "conduct", "cp", "ts", "tl", "tk", "erodibility"]
    p["phases"] = []
    for _ in range(p["nphase"]):
        r.advance()
        v = r.read(len(names), [int] + [float] * (len(names) - 1))
        p["phases"].append(dict(zip(names, v)))
    r.advance(); p["irphase"] = r.read(1, int)[0]
    r.advance()
    if p["irphase"] > 0:
        p["phasefile"] = r.read(1, str)[0]
    else:
        r.skip_record()
    r.advance(); p["inhom"] = r.read(1, int)[0]
    if p["inhom"] > abi.MAXINH:
        raise InpError("Read_params: Increase arrays for inhomogenities (reference: STOP 26)")
    p["inhoms"] = []
    for _ in range(p["inhom"]):
        r.advance()
        v = r.read(7, [int] * 6 + [float])
        p["inhoms"].append(dict(zip(["ix1", "ix2", "iy1", "iy2", "inphase", "igeom", "xinitaps"], v)))
    r.advance(); p["ten_off"] = r.read(1)[0]
    r.advance(); p["tau_heal"] = r.read(1)[0]
    r.advance(); p["v_min"], p["v_max"], p["ivis_shape"], p["efoldc"] = r.read(4, [float, float, int, float])
    r.advance()
    (p["itype_melting"], p["nelem_serp"], p["nelem_dike"], p["prod_magma"], p["prod_magma2"],
     p["rho_magma"]) = r.read(6, [int, int, int, float, float, float])
    r.advance()
    (p["angle_mzone"], p["fmagma_max"], p["ratio_mantle_mzone"],
     p["ratio_crust_mzone"]) = r.read(4)
    r.advance(); p["latent_heat_magma"], p["lambda_freeze"], p["lambda_freeze_tdep"] = r.read(3)
    r.advance(); p["weaken_ratio_plastic"], wrv = r.read(2)
    p["weaken_ratio_viscous"] = math.log(wrv)
    # remeshing
    r.advance(); p["ny_rem"], p["mode_rem"], p["ntest_rem"], p["angle_rem"] = r.read(4, [int, int, int, float])
    if p["mode_rem"] not in (1, 11, 3, 4):
        raise InpError("Illegal remeshing mode! Allowable - 1, 3, 4or 11")
    r.advance(); p["dx_rem"] = r.read(1)[0]
    r.advance(); p["topo_kappa"], p["fac_kappa"] = r.read(2)
    # process control
    r.advance(); p["idt_scale"] = r.read(1, int)[0]
    r.advance(); p["dt_scale"], p["strain_inert"] = r.read(2)
    r.advance(); p["i_rey"], p["xReyn"] = r.read(2, [int, float])
    r.advance(); p["ifreq_rmasses"] = r.read(1, int)[0]
    r.advance(); p["ifreq_imasses"] = r.read(1, int)[0]
    r.advance(); p["ifreq_visc"] = r.read(1, int)[0]
    r.advance(); p["ifreq_avgsr"] = r.read(1, int)[0]
    nt = p["ntest_rem"]
    if (nt % p["ifreq_rmasses"] + nt % p["ifreq_imasses"] + nt % p["ifreq_visc"]
            + nt % p["ifreq_avgsr"]) != 0:
        raise InpError("ntest_rem must be multiples of process frequency")
    r.advance(); p["frac"], p["fracm"] = r.read(2)
    r.advance(); p["movegrid"], p["ndim"] = r.read(2, int)
    r.advance(); p["demf"], p["mix_strain"], p["mix_stress"] = r.read(3, [float, int, int])
    # output
    r.advance(); p["time_max"] = r.read(1)[0] * 1000 * SEC_YEAR
    r.advance(); p["dtout_screen"] = r.read(1)[0] * 1000 * SEC_YEAR
    r.advance(); p["dtout_file"] = r.read(1)[0] * 1000 * SEC_YEAR
    r.advance(); p["io_flags"] = r.read(19, int)
    r.advance(); p["lastout"] = r.read(1, int)[0]
    r.advance(); p["dtsave_file"] = r.read(1)[0] * 1000 * SEC_YEAR
    r.advance(); p["lastsave"] = r.read(1, int)[0]
    return p


def read_inp(path):
    with open(path) as f:
        return parse_inp(f.read())


def to_gf_params(p, vbc, ivis_present, pisos=None):
    """Fill the ABI parameter block from a parsed `.inp` dict plus the values the
    reference derives during set-up (vbc from init_bc/vbcal, ivis_present from
    check_visc_rheol, pisos from init_stress when nyhydro == 2)."""
    g = abi.GfParams()
    for k in ("nx", "nz", "nphase", "mix_strain", "mix_stress", "ndim", "movegrid", "itherm",
              "istress_therm", "ishearh", "itemp_bc", "itype_melting", "nystressbc", "nydrsides",
              "nyhydro", "i_prestress", "idt_scale", "i_rey", "iphsub", "ifreq_rmasses",
              "ifreq_imasses", "ifreq_visc", "ifreq_avgsr", "inhom"):
        setattr(g, k, int(p[k]))
    g.ivis_present = int(ivis_present)
    for k in ("t_top", "t_bot", "bot_bc", "g", "demf", "ten_off", "tau_heal", "v_min", "v_max",
              "frac", "fracm", "strain_inert", "dt_scale", "xReyn", "rxbo", "rzbo", "drosub",
              "fmagma_max", "weaken_ratio_plastic", "weaken_ratio_viscous", "rho_magma"):
        setattr(g, k, float(p[k]))
    g.vbc = float(vbc)
    g.topo_kappa = float(p.get("topo_kappa", 0.0))
    g.flexure_base = int(any(b["nbc"] == 200 and b["nofside"] == 2 for b in p["bcs"]))
    g.nelem_serp = int(p["nelem_serp"])
    g.pisos = float(p["pisos"] if pisos is None else pisos)
    if p["nphase"] > abi.MAXPH:
        raise InpError("nphase exceeds maxph")
    for k, ph in enumerate(p["phases"]):
        g.irheol[k] = int(ph["irheol"])
        for name in abi.PHASE_TABLES:
            getattr(g, name)[k] = float(ph[name])
        g.erodibility[k] = float(ph["erodibility"])
    for k, ih in enumerate(p["inhoms"]):
        for name in ("ix1", "ix2", "iy1", "iy2", "igeom"):
            getattr(g, name)[k] = int(ih[name])
    return g


def read_params_txt(path):
    """Parameter block dumped by the reference harness (oracle/ref_harness.f90),
    used to check this reader against the reference's own read_params."""
    g = abi.GfParams()
    iflags = ["mix_strain", "mix_stress", "ndim", "movegrid", "itherm", "istress_therm",
              "ishearh", "itemp_bc", "itype_melting", "nystressbc", "nydrsides", "nyhydro",
              "i_prestress", "idt_scale", "i_rey", "ivis_present", "iphsub", "ifreq_rmasses",
              "ifreq_imasses", "ifreq_visc", "ifreq_avgsr", "inhom"]
    reals = ["t_top", "t_bot", "bot_bc", "g", "demf", "ten_off", "tau_heal", "v_min", "v_max",
             "frac", "fracm", "strain_inert", "dt_scale", "xReyn", "vbc", "rxbo", "rzbo", "pisos",
             "drosub", "fmagma_max", "weaken_ratio_plastic", "weaken_ratio_viscous", "rho_magma",
             "extra_pres"]
    ninh = 0
    with open(path) as f:
        for line in f:
            t = line.split()
            if t[0] == "nx_nz_nphase":
                g.nx, g.nz, g.nphase = (int(v) for v in t[1:4])
            elif t[0] == "flags":
                for k, v in zip(iflags, t[1:]):
                    setattr(g, k, int(v))
            elif t[0] == "flags2":
                g.nelem_serp = int(t[1])
                if len(t) > 2:
                    g.nopbmax = int(t[2])
            elif t[0] == "reals":
                for k, v in zip(reals, t[1:]):
                    if hasattr(g, k):
                        setattr(g, k, float(v))
            elif t[0] == "phase":
                k = int(t[1]) - 1
                g.irheol[k] = int(t[2])
                for name, v in zip(abi.PHASE_TABLES, t[3:]):
                    getattr(g, name)[k] = float(v)
            elif t[0] == "topo":
                g.topo_kappa = float(t[1])
                for k, v in enumerate(t[2:]):
                    g.erodibility[k] = float(v)
            elif t[0] == "inhom":
                for name, v in zip(("ix1", "ix2", "iy1", "iy2"), t[1:5]):
                    getattr(g, name)[ninh] = int(v)
                g.igeom[ninh] = int(t[6])
                ninh += 1
    return g
