"""Input decks in the reference's format (src/deck.cpp:14-186, src/sim_config.cpp:69-343): blocks `name { key =
value }`, comments from '#' outside quotes, several assignments on a line separated by ';', repeated `laser`, `ion`
and `bunch` blocks named laser2, ion2, ...; values with physical units; command-line style overrides
`block.key=value`.

`read_deck(path)` returns the engine's run configuration (g3d_amd.Deck) for the blocks and keys the march uses:
grid (Fourier or table y, uniform or table zeta), laser (plane_wave, gaussian), plasma (uniform or parabolic, the z
ramp), ion (by element or by mass, ADK), Gaussian electron bunches (bunch, bunch2, ... up to 4), numerics and the
native output settings.  Decks that need what the engine does not have (expressions and the block `constants`, file
lasers, profile and envelope functions, density functions, flat-top or positron bunches, undulators and magnets,
openPMD output, diag groups) are refused with an error that names the key; so are unknown keys, as in the reference
(Deck::check_unused, src/deck.cpp:143-148).
"""
import math
import re

import g3d_amd

# units of src/deck.cpp:158-168: unit -> (kind, factor to SI); energies in eV
UNITS = {
    "nm": ("length", 1e-9), "um": ("length", 1e-6), "mm": ("length", 1e-3), "cm": ("length", 1e-2),
    "m": ("length", 1.0), "fs": ("time", 1e-15), "ps": ("time", 1e-12), "ns": ("time", 1e-9),
    "pC": ("charge", 1e-12), "nC": ("charge", 1e-9), "cm^-3": ("density", 1e6), "m^-3": ("density", 1.0),
    "rad": ("angle", 1.0), "mrad": ("angle", 1e-3), "deg": ("angle", math.pi / 180),
    "u": ("mass", 1.66053906660e-27), "m_p": ("mass", 1.67262192369e-27), "m_e": ("mass", 9.1093837015e-31),
    "kg": ("mass", 1.0), "eV": ("energy", 1.0), "keV": ("energy", 1e3), "MeV": ("energy", 1e6),
    "GeV": ("energy", 1e9), "T": ("field", 1.0), "mT": ("field", 1e-3), "T/m": ("gradient", 1.0)}
ME_SI, AMU_SI, MC2_EV = 9.1093837015e-31, 1.66053906660e-27, 0.51099895e6

# the elements up to Si: nuclear charge, standard atomic mass [u] and ionization energies [eV] of the charge states
# 0 ... Z-1 (NIST Atomic Spectra Database values, the same as the reference's table src/atomic_data.cpp:28-52)
ELEMENTS = {
    "H": (1, 1.008, [13.598434599702]),
    "D": (1, 2.014, [13.598434599702]),
    "He": (2, 4.0026, [24.587389011, 54.4177655282]),
    "Li": (3, 6.94, [5.391714996, 75.6400970, 122.45435913]),
    "Be": (4, 9.0122, [9.322699, 18.21115, 153.896205, 217.71858459]),
    "B": (5, 10.81, [8.298019, 25.15483, 37.93059, 259.374379, 340.2260225]),
    "C": (6, 12.011, [11.2602880, 24.383143, 47.88778, 64.49352, 392.09056, 489.99320779]),
    "N": (7, 14.007, [14.53413, 29.60125, 47.4453, 77.4735, 97.8901, 552.06741, 667.0461377]),
    "O": (8, 15.999, [13.618055, 35.12112, 54.93554, 77.41350, 113.8990, 138.1189, 739.32697, 871.4099138]),
    "F": (9, 18.998, [17.42282, 34.97081, 62.70798, 87.175, 114.249, 157.16311, 185.1868, 953.8983, 1103.1175302]),
    "Ne": (10, 20.180, [21.564541, 40.96297, 63.4233, 97.1900, 126.247, 157.934, 207.271, 239.0970, 1195.8082,
                        1362.199256]),
    "Na": (11, 22.990, [5.13907696, 47.28636, 71.6200, 98.936, 138.404, 172.23, 208.504, 264.192, 299.856,
                        1465.0992, 1648.702285]),
    "Mg": (12, 24.305, [7.646236, 15.035271, 80.1436, 109.2654, 141.33, 186.76, 225.02, 265.924, 327.99, 367.489,
                        1761.8049, 1962.663889]),
    "Al": (13, 26.982, [5.985769, 18.82855, 28.447642, 119.9924, 153.8252, 190.49, 241.76, 284.64, 330.21, 398.65,
                        442.005, 2085.97693, 2304.140359]),
    "Si": (14, 28.085, [8.15168, 16.34585, 33.49300, 45.14179, 166.767, 205.279, 246.57, 303.59, 351.28, 401.38,
                        476.273, 523.415, 2437.65805, 2673.177958]),
}
HEAVY = ("Ar", "Cu", "Kr", "Ag", "Xe", "Au")              # in the reference's table; more levels than the engine holds
REPEATED = {"laser", "ion", "bunch", "undulator", "quadrupole", "plasma_lens", "solenoid", "chicane"}
_NUM = re.compile(r"\s*([+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)")
_STRTOD = re.compile(r"\s*[+-]?(?:\d|\.\d|inf|nan)", re.IGNORECASE)


class DeckError(ValueError):
    pass


def parse_blocks(text, name="(deck)"):
    """the blocks of a deck text, {block: {key: value}} (src/deck.cpp:14-60)"""
    blocks, block = {}, ""
    for ln, line in enumerate(text.split("\n"), 1):
        q = False
        for c, ch in enumerate(line):
            if ch == '"':
                q = not q
            elif ch == "#" and not q:
                line = line[:c]
                break
        line = line.strip(" \t\r")
        if not line:
            continue
        if not block:
            br = line.find("{")
            if br < 0:
                raise DeckError("%s:%d: expected 'block {'" % (name, ln))
            block = line[:br].strip(" \t\r")
            if block in blocks:
                if block not in REPEATED:
                    raise DeckError("%s:%d: block '%s' given twice" % (name, ln, block))
                k = 2
                while block + str(k) in blocks:
                    k += 1
                block += str(k)
            blocks[block] = {}
            line = line[br + 1:].strip(" \t\r")
            if not line:
                continue
        close = line.endswith("}")
        if close:
            line = line[:-1].strip(" \t\r")
        parts, q, st = [], False, 0
        for c in range(len(line) + 1):
            if c < len(line) and line[c] == '"':
                q = not q
            if c == len(line) or (line[c] == ";" and not q):
                parts.append(line[st:c].strip(" \t\r"))
                st = c + 1
        for part in parts:
            if not part:
                continue
            eq = part.find("=")
            if eq < 0:
                raise DeckError("%s:%d: expected 'key = value'" % (name, ln))
            blocks[block][part[:eq].strip(" \t\r")] = part[eq + 1:].strip(" \t\r")
        if close:
            block = ""
    if block:
        raise DeckError("%s: block '%s' is not closed" % (name, block))
    return blocks


def parse_with_unit(text, kind, where=""):
    """a number with an optional unit of the given kind ('length', 'time', ...; None: dimensionless) in SI; a bare
    angle is in radians (src/deck.cpp:155-180)"""
    m = _NUM.match(text)
    if not m:
        raise DeckError("%s: number expected, got '%s' (expressions are not supported)" % (where, text))
    v = float(m.group(1))
    rest = text[m.end():].split()
    if not rest:
        if kind is None or kind == "angle":
            return v
        raise DeckError("%s: a unit is required" % where)
    if kind is None:
        raise DeckError("%s: dimensionless value expected" % where)
    if rest[0] not in UNITS:
        raise DeckError("%s: unknown unit '%s'" % (where, rest[0]))
    uk, f = UNITS[rest[0]]
    if uk != kind:
        raise DeckError("%s: unit '%s' is of the wrong kind" % (where, rest[0]))
    return v * f


class _Reader:
    """the Deck accessors of src/deck.cpp:84-148 over parsed blocks, recording the keys read"""

    def __init__(self, blocks, name):
        self.b, self.name, self.used = blocks, name, set()

    def where(self, b, k):
        return "%s: %s.%s" % (self.name, b, k)

    def has(self, b, k):
        return b in self.b and k in self.b[b]

    def has_block(self, b):
        return b in self.b

    def str(self, b, k, default=None):
        if not self.has(b, k):
            if default is None:
                raise DeckError("%s: missing key %s.%s" % (self.name, b, k))
            return default
        self.used.add((b, k))
        v = self.b[b][k]
        return v[1:-1] if len(v) >= 2 and v[0] == '"' and v[-1] == '"' else v

    def num(self, b, k, kind, default=None):
        if default is not None and not self.has(b, k):
            return default
        t = self.str(b, k)
        try:
            return parse_with_unit(t, kind, self.where(b, k))
        except DeckError:
            # "number [unit]" with a known unit stays an error; the reference reads anything else as an
            # expression (src/deck.cpp:96-107)
            m = _NUM.match(t)
            rest = t[m.end():].split() if m else None
            if m and len(rest) <= 1 and (not rest or rest[0] in UNITS or rest[0] in ("k0^-1", "n_c")):
                raise
            raise DeckError("%s: '%s' is an expression; expressions are not supported by this engine"
                            % (self.where(b, k), t))

    def integer(self, b, k, default):
        if not self.has(b, k):
            return default
        v = self.num(b, k, None)
        if v != math.floor(v):
            raise DeckError(self.where(b, k) + ": integer expected")
        return int(v)

    def list(self, b, k, kind):
        tok = self.str(b, k).replace(",", " ").split()
        unit = ""
        if tok and not _STRTOD.match(tok[-1]):
            unit = tok.pop()
        if not tok:
            raise DeckError(self.where(b, k) + ": list of numbers expected")
        return [parse_with_unit(w + " " + unit if unit else w, kind, self.where(b, k)) for w in tok]

    def flag(self, b, k, default):
        if not self.has(b, k):
            return default
        v = self.str(b, k)
        if v in ("true", "on", "yes", "1"):
            return True
        if v in ("false", "off", "no", "0"):
            return False
        raise DeckError(self.where(b, k) + ": true or false expected")

    def refuse(self, b, k, why="is not supported by this engine"):
        if self.has(b, k):
            raise DeckError("%s %s" % (self.where(b, k), why))

    def check_unused(self):
        for b, kv in self.b.items():
            for k in kv:
                if (b, k) not in self.used:
                    raise DeckError(self.where(b, k) + ": unknown key")


def read_deck(path, overrides=()):
    """the engine's configuration of a deck file and command-line overrides `block.key=value`
    (src/main.cpp:30-42, src/sim_config.cpp:69-343)"""
    with open(path) as f:
        return deck_from_text(f.read(), overrides, str(path))


def deck_from_text(text, overrides=(), name="(deck)"):
    blocks = parse_blocks(text, name)
    for a in overrides:                                   # src/deck.cpp:68-76
        eq = a.find("=")
        dot = a.rfind(".", 0, eq) if eq >= 0 else -1
        if eq < 0 or dot < 0:
            raise DeckError("override '%s' must read block.key=value" % a)
        blocks.setdefault(a[:dot].strip(" \t\r"), {})[a[dot + 1:eq].strip(" \t\r")] = a[eq + 1:].strip(" \t\r")
    d = _Reader(blocks, name)
    for b in blocks:
        base = b.rstrip("0123456789")
        if b in ("constants", "diag") or b.startswith("diag.") or base in (
                "undulator", "quadrupole", "plasma_lens", "solenoid", "chicane"):
            raise DeckError("%s: block '%s' is not supported by this engine" % (name, b))
    lam = d.num("grid", "lambda0", "length",
                d.num("laser", "wavelength", "length") if d.has("laser", "wavelength") else 1e-6)
    deck = g3d_amd.Deck(lam)
    c = deck.c
    Lu, nc = deck.Lu, deck.nc
    Tu = Lu / g3d_amd.C_SI
    L = lambda b, k: d.num(b, k, "length") / Lu
    Ld = lambda b, k, dflt: d.num(b, k, "length") / Lu if d.has(b, k) else dflt
    LL = lambda k: [v / Lu for v in d.list("grid", k, "length")]
    _grid(d, deck, L, Ld, LL)
    _lasers(d, deck, L, Ld, lam, Tu)
    _plasma(d, deck, L, Ld, Lu, nc)
    _ions(d, deck, nc)
    c.iterations = d.integer("numerics", "iterations", 2)
    c.vz_min = d.num("numerics", "vz_min", None, 0.8)
    c.vz_beam = d.num("numerics", "vz_beam", None, 0.9)
    c.trap_advance = d.num("numerics", "trap_advance", None, 1.0)
    pr = d.str("numerics", "predictor", "extrapolate")
    if pr not in ("extrapolate", "previous"):
        raise DeckError("numerics.predictor must be extrapolate or previous")
    c.predictor_extrapolate = int(pr == "extrapolate")
    if c.iterations < 1:
        raise DeckError("numerics.iterations must be >= 1")
    if not (0 < c.vz_min < c.vz_beam < 1):
        raise DeckError("numerics: need 0 < vz_min < vz_beam < 1")
    if not c.trap_advance >= 1:
        raise DeckError("numerics.trap_advance must be >= 1")
    if (c.plasma_on or c.n_glasers) and not c.self_fields:
        raise DeckError("numerics.self_fields = false is for test particles in the analytic laser only")
    _bunch(d, deck, L, Ld)
    deck.output = _output(d, Lu)
    c.freeze_bunch = int(d.flag("numerics", "freeze_bunch", False))
    c.start_relax = int(d.flag("numerics", "start_relax", True))
    bm = d.str("numerics", "box_mean", "keep")
    if bm not in ("keep", "remove"):
        raise DeckError("numerics.box_mean must be keep or remove")
    c.box_mean = int(bm == "keep")
    if not c.box_mean and not c.y_fourier:
        raise DeckError("numerics.box_mean = remove needs the Fourier y grid")
    d.check_unused()
    return deck


def _grid(d, deck, L, Ld, LL):
    """grid block and the numerics switches read with it (src/sim_config.cpp:84-113, 14-48)"""
    c = deck.c
    c.nx = d.integer("grid", "nx", 0)
    if c.nx < 4:
        raise DeckError("grid: nx must be >= 4")
    c.Lx = L("grid", "Lx")
    yg = d.str("grid", "y_grid", "fourier")
    if yg == "fourier":
        ny = d.integer("grid", "ny", 0)
        if ny < 4:
            raise DeckError("grid: ny must be >= 4")
        deck.y_uniform(ny, Ld("grid", "Ly", c.Lx))
    elif yg == "table":
        if d.has("grid", "ny") or d.has("grid", "Ly"):
            raise DeckError("grid: with y_grid = table, ny and Ly follow from y_points, dy_points and y_max")
        if d.flag("grid", "move_y", False):
            raise DeckError("grid: move_y needs y_grid = fourier (the free-hand y grid is fixed between walls)")
        deck.y_table(LL("y_points"), LL("dy_points"), L("grid", "y_max"))
    else:
        raise DeckError("grid.y_grid must be fourier or table")
    zg = d.str("grid", "zeta_grid", "uniform")
    if zg == "uniform":
        nz = d.integer("grid", "nzeta", 0)
        if nz < 4:
            raise DeckError("grid: nzeta must be >= 4")
        deck.zeta_uniform(nz, L("grid", "window_length"))
    elif zg == "table":
        if d.has("grid", "nzeta"):
            raise DeckError("grid: with zeta_grid = table, nzeta follows from the table and window_length")
        deck.zeta_table(LL("zeta_points"), LL("dzeta_points"), L("grid", "window_length"))
    else:
        raise DeckError("grid.zeta_grid must be uniform or table")
    c.dz = L("grid", "dz")
    c.z_end = L("grid", "z_end")
    c.dz_push = Ld("grid", "dz_push", c.dz)
    c.sponge = Ld("grid", "sponge", 0.1 * min(c.Lx, c.Ly))
    c.sponge_y = Ld("grid", "sponge_y", c.sponge)
    c.sponge_rate = d.num("grid", "sponge_rate", None, 0.1)
    c.move_x = int(d.flag("grid", "move_x", True))
    if c.y_fourier:
        c.move_y = int(d.flag("grid", "move_y", True))
    c.window_lag = d.integer("grid", "window_lag", 4)
    if c.window_lag < 1:
        raise DeckError("grid.window_lag must be >= 1")
    c.self_fields = int(d.flag("numerics", "self_fields", True))
    c.zeta_filter = int(d.flag("numerics", "zeta_filter", True))
    c.k_filter = int(d.flag("numerics", "k_filter", True))
    c.push_phase = d.num("numerics", "push_phase", "angle", 0.05)


def _lasers(d, deck, L, Ld, lam, Tu):
    """laser blocks: at most one analytic plane wave, Gaussian lasers on the grid (src/sim_config.cpp:115-199)"""
    c = deck.c
    k = 1
    while True:
        b = "laser" if k == 1 else "laser%d" % k
        k += 1
        if not d.has_block(b):
            if k == 2:
                continue
            break
        typ = d.str(b, "type", "plane_wave")
        if typ == "plane_wave":
            if c.laser_on or c.a0 != 0:
                raise DeckError(b + ": only one laser of type plane_wave")
            c.laser_on = int(d.flag(b, "enabled", True))
            c.a0 = d.num(b, "a0", None)
            c.kL = lam / d.num(b, "wavelength", "length", lam)
            c.theta = d.num(b, "angle", "angle", 0.0)
            c.tau = d.num(b, "tau_fwhm", "time") / Tu
            c.cep = d.num(b, "cep", "angle", 0.0)
            pol = d.str(b, "polarization", "s")
            if pol not in ("s", "p", "circular"):
                raise DeckError(b + ".polarization must be s, p or circular")
            c.laser_pol = {"s": 0, "p": 1, "circular": 2}[pol]
            c.start_fwhm = d.num(b, "start_fwhm", None, 3.5)
            c.stop_fwhm = d.num(b, "stop_fwhm", None, c.start_fwhm)
        elif typ == "gaussian":
            if not d.flag(b, "enabled", True):
                continue
            d.refuse(b, "profile_function")
            d.refuse(b, "envelope_function")
            if c.n_glasers >= 4:
                raise DeckError(b + ": at most 4 grid lasers")
            a0 = d.num(b, "a0", None)
            kL = lam / d.num(b, "wavelength", "length", lam)
            waist = L(b, "waist")
            tau = d.num(b, "tau_fwhm", "time") / Tu
            zc = L(b, "zeta_center")
            zf = Ld(b, "z_focus", 0.0)
            cep = d.num(b, "cep", "angle", 0.0)
            pol = d.str(b, "polarization", "linear")
            if pol not in ("linear", "circular"):
                raise DeckError(b + ".polarization must be linear or circular")
            pa = d.num(b, "pol_angle", "angle", 0.0)
            hel = d.integer(b, "helicity", 1)
            if hel not in (1, -1):
                raise DeckError(b + ".helicity must be 1 or -1")
            deck.add_grid_laser(a0, waist, tau, zc, z_focus=zf, kL=kL, cep=cep, pol_angle=pa)
            g = c.glasers[c.n_glasers - 1]
            g.circular, g.helicity = int(pol == "circular"), hel
        elif typ == "file":
            raise DeckError(b + ".type = file is not supported by this engine")
        else:
            raise DeckError(b + ".type must be plane_wave, gaussian or file")


def _plasma(d, deck, L, Ld, Lu, nc):
    """plasma block (src/sim_config.cpp:201-229)"""
    c = deck.c
    if not d.has_block("plasma"):
        return
    c.plasma_on = 1
    dn = d.str("plasma", "density")
    if len(dn) > 3 and dn.endswith("n_c"):
        c.n0 = parse_with_unit(dn[:-3], None, "plasma.density")
    else:
        c.n0 = d.num("plasma", "density", "density") / nc
    if not c.n0 > 0:
        raise DeckError("plasma.density must be positive")
    if d.has("plasma", "z_points") or d.has("plasma", "z_values"):
        pts = [v / Lu for v in d.list("plasma", "z_points", "length")]
        vals = d.list("plasma", "z_values", None)
        if len(pts) != len(vals) or not pts:
            raise DeckError("plasma: z_points and z_values need the same number of entries")
        if any(not pts[k] > pts[k - 1] for k in range(1, len(pts))):
            raise DeckError("plasma.z_points must increase")
        if len(pts) > 16:
            raise DeckError("plasma.z_points: at most 16 points in this engine")
        c.n_pz = len(pts)
        for i, (p, v) in enumerate(zip(pts, vals)):
            c.pz_points[i], c.pz_values[i] = p, v
    prof = d.str("plasma", "profile", "uniform")
    if prof == "parabolic":
        c.parabolic = 1
        c.channel_depth = d.num("plasma", "channel_depth", None)
        c.r_channel = L("plasma", "r_channel")
    elif prof != "uniform":
        raise DeckError("plasma.profile must be uniform or parabolic")
    c.r_max = Ld("plasma", "r_max", 1e300)
    c.ppc_x = d.integer("plasma", "ppc_x", 2)
    c.ppc_y = d.integer("plasma", "ppc_y", c.ppc_x)
    if c.ppc_x < 1 or c.ppc_y < 1:
        raise DeckError("plasma: ppc_x, ppc_y must be >= 1")
    c.smooth_length = Ld("plasma", "smooth_length", 0.0)
    d.refuse("plasma", "density_function")
    if c.smooth_length < 0:
        raise DeckError("plasma.smooth_length must be >= 0")


def _ions(d, deck, nc):
    """ion blocks and plasma.free_electrons (src/sim_ions.cpp:6-62)"""
    c = deck.c
    c.free_electrons = int(d.flag("plasma", "free_electrons", True))
    qsum, k = 0.0, 1
    while d.has_block("ion" if k == 1 else "ion%d" % k):
        b = "ion" if k == 1 else "ion%d" % k
        if not c.plasma_on:
            raise DeckError(b + ": ion blocks need a plasma block")
        mode = d.str(b, "ionization", "none")
        if mode not in ("none", "adk"):
            raise DeckError(b + ".ionization must be none or adk")
        ionization = mode == "adk"
        el, ms = d.has(b, "element"), d.has(b, "mass")
        if el == ms:
            raise DeckError(b + ": give either element or mass")
        Znuc, ip, sym = 0, [], ""
        if el:
            given = d.str(b, "element")                  # case-insensitive, the canonical symbol kept
            sym = next((e for e in list(ELEMENTS) + list(HEAVY) if e.lower() == given.lower()), given)
            if sym in HEAVY:
                raise DeckError("%s.element: %s has more than the 16 ionization levels this engine holds" % (b, sym))
            if sym not in ELEMENTS:
                raise DeckError("%s.element: unknown element '%s'" % (b, sym))
            Znuc, mu, ip = ELEMENTS[sym]
            mass = mu * AMU_SI / ME_SI
            if not ionization and Znuc > 1 and not d.has(b, "charge_state"):
                raise DeckError("%s: charge_state is required for %s" % (b, sym))
        else:
            mass = d.num(b, "mass", "mass") / ME_SI
            if not mass > 0:
                raise DeckError(b + ".mass must be positive")
        Z0 = d.integer(b, "charge_state", 0 if ionization else 1)
        if d.has(b, "ionization_potentials"):
            ipl = d.list(b, "ionization_potentials", "energy")
            if el and len(ipl) != Znuc:
                raise DeckError("%s.ionization_potentials: give all %d levels of %s" % (b, Znuc, sym))
            ip = ipl
            if not el:
                Znuc = len(ip)
        if Znuc == 0:
            Znuc = max(Z0, 1)
        if Z0 < 0 or Z0 > Znuc:
            raise DeckError("%s.charge_state must be 0 ... %d" % (b, Znuc))
        if ionization:
            if not ip:
                raise DeckError(b + ": ionization of an ion given by mass needs ionization_potentials")
            Zmax = d.integer(b, "max_charge_state", Znuc)
            if Zmax <= Z0 or Zmax > Znuc:
                raise DeckError("%s.max_charge_state must be > charge_state and <= %d" % (b, Znuc))
        else:
            if Z0 < 1:
                raise DeckError(b + ".charge_state must be >= 1 without ionization")
            if d.has(b, "max_charge_state"):
                raise DeckError(b + ".max_charge_state needs ionization = adk")
            Zmax = Z0
        if d.has(b, "fraction") and d.has(b, "density"):
            raise DeckError(b + ": give either fraction or density")
        if d.has(b, "density"):
            atoms = d.num(b, "density", "density") / nc / c.n0
        elif Z0 >= 1:
            atoms = d.num(b, "fraction", None, 1.0) / Z0
        else:
            raise DeckError(b + ": neutral atoms (charge_state = 0) need density")
        if not atoms > 0:
            raise DeckError(b + ": fraction (density) must be positive")
        if not c.free_electrons and Z0 != 0:
            raise DeckError(b + ": with plasma.free_electrons = false all ions start neutral (charge_state = 0)")
        mobile = d.flag(b, "mobile", True)
        nm = d.str(b, "name", (sym + ("%d+" % Z0 if Z0 > 1 else "")) if el else "ion%d" % k)
        if len(nm.encode()) > 15:
            raise DeckError(b + ".name: at most 15 characters in this engine")
        if c.n_ions >= 4:
            raise DeckError(b + ": at most 4 ion species in this engine")
        if len(ip) > 16:
            raise DeckError(b + ": at most 16 ionization levels in this engine")
        qsum += Z0 * atoms
        deck.add_ion(mass, Z0, Zmax, atoms=atoms, mobile=mobile, ionization=ionization, ip_eV=ip, name=nm)
        k += 1
    if qsum > 1 + 1e-12:
        raise DeckError("ion blocks: the ion charge (sum of Z0 n_atom) exceeds the free-electron density")


def _bunch(d, deck, L, Ld):
    """the bunch blocks bunch, bunch2, ...: Gaussian electron bunches with random shot noise (src/sim_bunches.cpp:
    6-82); the first sets the single-bunch fields, the others are the further bunches (species -1, -2, -3)"""
    c = deck.c
    lw = float(deck.zeta[-1])
    if not d.has_block("bunch"):
        if d.has_block("bunch2"):
            raise DeckError("bunch2 without a block bunch")
        c.nmacro, c.charge, c.gamma, c.zeta_center = 0, 0.0, 1.0, 0.75 * lw
        return
    names = []
    k = 1
    while d.has_block("bunch" if k == 1 else "bunch%d" % k):
        b = "bunch" if k == 1 else "bunch%d" % k
        name = d.str(b, "name", b)
        if k == 1 and name != "bunch":
            raise DeckError("bunch.name: this engine names its first bunch 'bunch'")
        if name in names:
            raise DeckError("%s.name: '%s' given twice" % (b, name))
        names.append(name)
        for key in ("file", "gamma_function", "ux_function", "uy_function", "profile_z", "profile_r",
                    "noise_wavelength", "beamlet"):
            d.refuse(b, key)
        if d.str(b, "species", "electron") != "electron":
            raise DeckError(b + ".species: this engine loads electron bunches")
        charge = abs(d.num(b, "charge", "charge"))
        if d.has(b, "energy") + d.has(b, "gamma") != 1:
            raise DeckError(b + ": give one of energy, gamma")
        gamma = d.num(b, "gamma", None) if d.has(b, "gamma") else 1 + d.num(b, "energy", "energy") / MC2_EV
        if not gamma > 1:
            raise DeckError(b + ": gamma must be > 1")
        if d.str(b, "distribution_r", "gaussian") != "gaussian":
            raise DeckError(b + ".distribution_r: this engine loads Gaussian bunches")
        sigx = L(b, "sigma_x")
        sigy = Ld(b, "sigma_y", sigx)
        if d.str(b, "distribution_z", "gaussian") != "gaussian":
            raise DeckError(b + ".distribution_z: this engine loads Gaussian bunches")
        sigz = L(b, "sigma_z")
        emitx = Ld(b, "emittance_x", 0.0)
        emity = Ld(b, "emittance_y", emitx)
        espread = d.num(b, "energy_spread", None, 0.0)
        zc = Ld(b, "zeta_center", 0.75 * lw)
        if Ld(b, "x0", 0.0) != 0 or Ld(b, "y0", 0.0) != 0:
            raise DeckError(b + ": x0 and y0 other than 0 are not supported by this engine")
        nmacro = d.integer(b, "macroparticles", 100000)
        if d.str(b, "shot_noise", "random") != "random":
            raise DeckError(b + ".shot_noise: this engine draws random shot noise")
        seed = d.integer(b, "seed", k)
        if k == 1:
            c.charge, c.gamma, c.sigx, c.sigy, c.sigz = charge, gamma, sigx, sigy, sigz
            c.emitx, c.emity, c.espread, c.zeta_center, c.nmacro, c.seed = emitx, emity, espread, zc, nmacro, seed
        else:
            if k > 4:
                raise DeckError(b + ": at most 4 bunches in this engine")
            if len(name.encode()) > 15:
                raise DeckError(b + ".name: at most 15 characters in this engine")
            deck.add_bunch(name, charge, gamma, sigx, sigy, sigz, zc, nmacro, seed, emitx=emitx, emity=emity,
                           espread=espread)
        k += 1


def _output(d, Lu):
    """the main output group (src/sim_config.cpp:247-330): what the driver writes and when"""
    b = "output"
    out = {"dir": d.str(b, "dir", "out"), "every": d.integer(b, "every", 10),
           "particles": d.flag(b, "particles", True), "frames": d.flag(b, "frames", False),
           "slice_length": d.num(b, "slice_length", "length", 0.0) / Lu,
           "checkpoint_every": d.integer(b, "checkpoint_every", 0),
           "walltime_hours": d.num(b, "walltime_hours", None, 0.0)}
    if out["slice_length"] < 0:
        raise DeckError("output.slice_length must be >= 0")
    if out["checkpoint_every"] < 0 or out["walltime_hours"] < 0:
        raise DeckError("output: checkpoint_every and walltime_hours must be >= 0")
    out["beam_every"] = d.integer(b, "beam_every", 0)          # 0: no beam files from the main group
    if out["every"] < 0 or out["beam_every"] < 0:
        raise DeckError(b + ": every and beam_every must be >= 0")
    names = ["ex", "ey", "ez", "bx", "by", "bz", "ne", "ni"]
    fs = d.str(b, "fields", "all")
    if fs in ("all", "none"):
        out["fields"] = 255 if fs == "all" else 0
    else:
        out["fields"] = 0
        for w in fs.replace(",", " ").split():
            if w not in names:
                raise DeckError("%s.fields: unknown component '%s' (ex ey ez bx by bz ne ni, all, none)" % (b, w))
            out["fields"] |= 1 << names.index(w)
    bs = d.str(b, "beams", "all")
    if bs in ("all", "none"):
        out["beams"] = 7 if bs == "all" else 0
    else:
        out["beams"] = 0
        for w in bs.replace(",", " ").split():
            if w not in ("bunch", "trapped", "injected"):
                raise DeckError("%s.beams: unknown beam '%s' (bunch trapped injected, all, none)" % (b, w))
            out["beams"] |= {"bunch": 1, "trapped": 2, "injected": 4}[w]
    out["axis"] = d.flag(b, "axis", False)
    if d.str(b, "format", "native") != "native":
        raise DeckError("output.format: this engine writes the native format")
    for k in ("xi_stride", "particle_stride"):
        if d.integer(b, k, 1) != 1:
            raise DeckError("output.%s: this engine writes every slice and particle" % k)
    for k in ("openpmd_file", "openpmd_grid", "openpmd_dy", "openpmd_dzeta", "hdf5_single", "hdf5_compression"):
        d.refuse(b, k, "is for openPMD output, which this engine does not write")
    return out
