"""Cross-section tables for the null-collision MCC (host side, init time).

Restates the reference's table build so the GPU receives exactly the
MCCBundle arrays the Fortran code would hold:

* namelist input format  &MRList Model, NReaction / &SRList SR(k)%Reaction,
  SR(k)%EnergySigma        (R/src/mc/MCCSigma.f90:46-65, input/gas/lxcat/CF4/*.txt)
* SigmaNormalizedInit     (R/src/mc/MCCSigma.f90:87-129)
* SigmaNormalizedUpdate   (R/src/mc/MCCSigma.f90:131-190), including its use
  of dble(i) (not i-1) in the interpolation weight and the replacement of
  energies below 10**Emin by Emin itself
* SigmaNormalizedUpdateReactive (R/src/mc/MCCSigma.f90:192-220)
* ProbilityNonReactive / ProbilityReactive (R/src/mc/MCCInit.F90:125-170, 272-309)
* locate                  (R/src/base/numerical/Numrical.F90:112-145)
* gas density p*mTorrtoPa/(kB*T)*ratio (R/src/base/SpecyOne.F90:136) and the
  hard-coded polarizability 14*a0**3 (SpecyOne.F90:193)
"""
import math
import re
from dataclasses import dataclass, field
from typing import List

import numpy as np

# R/src/base/Constants.F90
PI = 3.141592653589793238
A0 = 5.2918e-11
KB = 1.3807e-23
ELECTRON_CHARGE = 1.6022e-19
ELECTRON_MASS = 9.1095e-31
ATOMIC_MASS = 1.660539040e-27
EPSILON0 = 8.8542e-12
JTOEV = 1.6022e-19
MTORR_TO_PA = 0.13332

# R/src/mc/MCCSigma.f90:10-14
ELECTRON_EMIN, ELECTRON_EMAX = -3.0, 3.0
ION_EMIN, ION_EMAX = -1.0, 3.0
NSIGMA_ELECTRON, NSIGMA_ION = 10001, 1001
NSIGMA_RAW_MAX = 500


@dataclass
class Reaction:
    reactant: int
    rtype: int
    resultant: int
    threshold: float
    energy: np.ndarray = None   # raw E [eV] (Model 1/2 only)
    sigma: np.ndarray = None    # raw sigma [m^2]


@dataclass
class SigmaFile:
    model: int
    reactions: List[Reaction] = field(default_factory=list)


@dataclass
class MCCBundle:
    """Device-ready table for one (species, gas): Type MCCBundle, R/src/mc/TypeMCC.f90:23-33."""
    model: int
    nreaction: int = 0
    nsigma: int = 0
    emin: float = 0.0
    einterval: float = 0.0
    emax: float = 0.0            # 10**EnergyMax
    collision_ratio: float = 0.0
    sigma_max: float = 0.0
    rtype: np.ndarray = None
    resultant: np.ndarray = None
    threshold: np.ndarray = None
    prob: np.ndarray = None      # shape (nsigma, nreaction): Probility(j,i) -> prob[i-1, j-1]
    gas_mass: float = 0.0
    gas_temperature: float = 0.0
    beta_max: float = 0.0


# ----------------------------------------------------------- namelist I/O
def _fortran_float(tok):
    return float(tok.replace("d", "e").replace("D", "e"))


def parse_sigma_file(text):
    """Parse the &MRList/&SRList namelist format of input/gas/lxcat/<gas>/<sp><gas>.txt."""
    body = "\n".join(line.split("!", 1)[0] for line in text.splitlines())
    m = re.search(r"&MRList(.*?)/", body, re.S | re.I)
    if not m:
        raise ValueError("no &MRList group")
    mr = m.group(1)
    model = int(re.search(r"Model\s*=\s*([-\d]+)", mr, re.I).group(1))
    nreac = int(re.search(r"NReaction\s*=\s*(\d+)", mr, re.I).group(1))
    rest = body[m.end():]
    m2 = re.search(r"&SRList(.*?)(?:^\s*/\s*$|/\s*\Z)", rest, re.S | re.I | re.M)
    if not m2:
        raise ValueError("no &SRList group")
    sr = m2.group(1)
    out = SigmaFile(model=model)
    if re.search(r"SR\s*%\s*Reaction\s*=", sr, re.I):       # array form: SR%Reaction = r,t,res,th ...
        toks = re.split(r"[,\s]+", sr.split("=", 1)[1].strip())
        toks = [t for t in toks if t]
        for k in range(nreac):
            a, b, c, d = toks[4 * k:4 * k + 4]
            out.reactions.append(Reaction(int(a), int(b), int(c), _fortran_float(d)))
        return out
    parts = re.split(r"SR\((\d+)\)\s*%\s*(Reaction|EnergySigma)\s*=", sr, flags=re.I)
    reac = {}
    es = {}
    for k in range(1, len(parts), 3):
        idx, key, val = int(parts[k]), parts[k + 1].lower(), parts[k + 2]
        nums = [t for t in re.split(r"[,\s]+", val.strip()) if t]
        if key == "reaction":
            reac[idx] = Reaction(int(nums[0]), int(nums[1]), int(nums[2]), _fortran_float(nums[3]))
        else:
            v = np.array([_fortran_float(t) for t in nums], dtype=np.float64)
            es[idx] = v.reshape(-1, 2)
    for k in range(1, nreac + 1):
        r = reac[k]
        if k in es:
            r.energy = es[k][:, 0].copy()
            r.sigma = es[k][:, 1].copy()
        out.reactions.append(r)
    return out


def write_sigma_file(model, reactions):
    """Write the namelist format parse_sigma_file reads (Model 1/2 with tables, 3 without)."""
    lines = ["&MRList", f"Model     =     {model},", f"NReaction  =    {len(reactions)},", "/", "", "&SRList"]
    if model == 3:
        lines.append("SR%Reaction= ")
        for r in reactions:
            lines.append(f"{r.reactant},{r.rtype},{r.resultant},{float(r.threshold)!r}")
    else:
        for k, r in enumerate(reactions, 1):
            lines.append(f"SR({k})%Reaction = {r.reactant},{r.rtype},{r.resultant},{float(r.threshold)!r},")
            lines.append(f"SR({k})%EnergySigma =")
            for e, s in zip(r.energy, r.sigma):
                lines.append(f" {float(e)!r}\t{float(s)!r}")
    lines.append("/")
    return "\n".join(lines) + "\n"


# ------------------------------------------------------------- algorithms
def locate(xx, n, x):
    """Numerical Recipes locate (R/src/base/numerical/Numrical.F90:112-145); 1-based result."""
    jl, ju = 0, n + 1
    asc = xx[n - 1] >= xx[0]
    while ju - jl > 1:
        jm = (ju + jl) // 2
        if asc == (x >= xx[jm - 1]):
            jl = jm
        else:
            ju = jm
    if x == xx[0]:
        return 1
    if x == xx[n - 1]:
        return n - 1
    return jl


def grid_for_model(model):
    """SigmaNormalizedInit (R/src/mc/MCCSigma.f90:87-129): (Emin, Eint, EnergyMax, NSigma)."""
    if model == 1:
        return ELECTRON_EMIN, (ELECTRON_EMAX - ELECTRON_EMIN) / float(NSIGMA_ELECTRON - 1), ELECTRON_EMAX, NSIGMA_ELECTRON
    if model == 2:
        return ION_EMIN, (ION_EMAX - ION_EMIN) / float(NSIGMA_ELECTRON - 1), ION_EMAX, NSIGMA_ELECTRON
    if model == 3:
        return ION_EMIN, (ION_EMAX - ION_EMIN) / float(NSIGMA_ION - 1), ION_EMAX, NSIGMA_ION
    raise ValueError(f"invalid model {model}")


def energy_grid(emin, eint, nsigma):
    i = np.arange(nsigma, dtype=np.float64)
    return np.power(10.0, emin + i * eint)


def sigma_normalize(reactions, emin, eint, nsigma):
    """SigmaNormalizedUpdate (R/src/mc/MCCSigma.f90:131-190).  Returns Sigma as (nsigma, nreaction)."""
    nr = len(reactions)
    out = np.zeros((nsigma, nr), dtype=np.float64)
    egrid = energy_grid(emin, eint, nsigma)
    ten_emin = 10.0 ** emin
    for j, r in enumerate(reactions):
        e = np.zeros(NSIGMA_RAW_MAX + 2)
        s = np.zeros(NSIGMA_RAW_MAX + 1)
        m = len(r.energy)
        if m > NSIGMA_RAW_MAX:
            raise ValueError("more raw points than NSigmaRawMax")
        e[:m] = r.energy
        s[:m] = r.sigma
        nraw = int(np.argmax(e[:NSIGMA_RAW_MAX])) + 1      # Maxloc: first maximum
        max_energy = e[nraw - 1]
        alfa = s[nraw - 1] * max_energy
        xx = e.copy()
        for i in range(nraw):
            xx[i] = emin if xx[i] < ten_emin else (math.log10(xx[i]) - emin) / eint
        th = r.threshold
        for i in range(1, nsigma + 1):
            energy = egrid[i - 1]
            if energy <= th:
                v = 0.0
            elif energy >= max_energy:
                v = alfa / energy
            else:
                k = locate(xx, nraw, float(i - 1))
                if k < 1:
                    k = 1
                le, ue = xx[k - 1], xx[k]
                s1 = (float(i) - le) / (ue - le)
                s2 = 1.0 - s1
                v = s2 * s[k - 1] + s1 * s[k]
            out[i - 1, j] = v
    return out


def sigma_normalize_reactive(thresholds, natom_species, natom_gas, emin, eint, nsigma):
    """SigmaNormalizedUpdateReactive (R/src/mc/MCCSigma.f90:192-220)."""
    ncomplex = natom_species + natom_gas
    S = (3.0 * ncomplex - 6.0) / 2.0 - 1.0
    egrid = energy_grid(emin, eint, nsigma)
    th = np.asarray(thresholds, dtype=np.float64)
    out = np.zeros((nsigma, len(th)))
    for i in range(nsigma):
        E = egrid[i]
        pt = np.array([math.pow(E - t, S) if E > t else 0.0 for t in th])
        psum = 0.0
        for v in pt:
            psum = psum + v
        out[i] = pt / psum
    return out


def probability_nonreactive(sigma, emin, eint, species_mass, gas_mass, gas_density):
    """ProbilityNonReactive (R/src/mc/MCCInit.F90:125-170).  Returns (prob, SigmaMax)."""
    nsigma, nr = sigma.shape
    miu = species_mass * gas_mass / (species_mass + gas_mass)
    egrid = energy_grid(emin, eint, nsigma)
    v = np.sqrt(2.0 * egrid * JTOEV / miu)
    p = sigma * v[:, None] * gas_density
    for j in range(1, nr):
        p[:, j] = p[:, j] + p[:, j - 1]
    mx = float(p.max())
    return p / mx, mx


def probability_reactive(sigma, species_mass, gas_mass, gas_density, beta_max, polarizability, dt):
    """ProbilityReactive (R/src/mc/MCCInit.F90:272-309).  Returns (prob, CollisionRatio)."""
    p = sigma.copy()
    miu = species_mass * gas_mass / (species_mass + gas_mass)
    ratio = math.sqrt(PI * polarizability * ELECTRON_CHARGE * ELECTRON_CHARGE / (EPSILON0 * miu)) \
        * beta_max * beta_max * gas_density * dt
    for j in range(1, p.shape[1]):
        p[:, j] = p[:, j] + p[:, j - 1]
    return p, ratio


def hard_sphere_sigma(species_radius, gas_radius):
    """SigmaMax of ProbilityHardShpere: 2 pi (r_g + r_s)/2 (r_g + r_s)/2
    (R/src/mc/MCCInit.F90:190), evaluated left to right as written."""
    return 2.0 * PI * 0.5 * (gas_radius + species_radius) * 0.5 * (gas_radius + species_radius)


def hard_sphere_bundle(species_index, species_mass, species_radius, gas, dt):
    """ProbilityHardShpere (R/src/mc/MCCInit.F90:172-270): the table an ion
    gets without a cross-section file, or against a gas it does not belong
    to (MCCBundleIonInit, MCCInit.F90:94-116).

    gas: dict(mass, temperature, density, radius, ns, index_start, beta_max).
    A gas with one ion species (GO%Ns == 1) gives two reactions, isotropic
    elastic 101 (resultant: the species) and charge exchange 103 (resultant:
    the gas's first ion, IndexStart + 1), with Probility = 0.5 and 1.0 times
    sigma v(E) n; any other gas gives the elastic one alone, 1.0 sigma v(E) n.
    Normalised by the maximum (SigmaMax), CollisionRatio = 1 - exp(-SigmaMax dt).

    The grid: NSigmaIon points with the ion energy range, log10 E from
    IonEnergyMin = -1 to IonEnergyMax = 3.  The reference takes EnergyMin and
    EnergyInterval from a SigmaNormalized that SigmaNormalization has just
    reset for the missing file (intent(out): EnergyInterval back to its
    default 0, EnergyMin undefined), which makes its later table lookup
    divide by zero; the NSigmaIon grid is the one the table's length
    implies."""
    emin, eint, emax_exp, nsigma = grid_for_model(3)
    sig = hard_sphere_sigma(species_radius, gas["radius"])
    if gas["ns"] == 1:
        base = [0.5, 1.0]
        rtype = [101, 103]
        res = [species_index, gas["index_start"] + 1]
    else:
        base = [1.0]
        rtype = [101]
        res = [species_index]
    miu = species_mass * gas["mass"] / (species_mass + gas["mass"])
    egrid = energy_grid(emin, eint, nsigma)
    v = np.sqrt(2.0 * egrid * JTOEV / miu)
    prob = np.empty((nsigma, len(base)))
    for j, b0 in enumerate(base):
        prob[:, j] = b0 * sig * v * gas["density"]
    mx = float(prob.max())
    prob = prob / mx
    return MCCBundle(model=2, nreaction=len(base), nsigma=nsigma, emin=emin, einterval=eint, emax=10.0 ** emax_exp,
                     collision_ratio=1.0 - math.exp(-mx * dt), sigma_max=mx,
                     rtype=np.array(rtype, dtype=np.int32), resultant=np.array(res, dtype=np.int32),
                     threshold=np.zeros(len(base)), prob=prob, gas_mass=gas["mass"],
                     gas_temperature=gas["temperature"], beta_max=gas.get("beta_max", 0.0))


def gas_density(pressure_mtorr, temperature, ratio=1.0):
    """GasInit: InitDensity = (p*mTorrtoPa)/(kB*T)*ratio (R/src/base/SpecyOne.F90:136)."""
    return (pressure_mtorr * MTORR_TO_PA) / (KB * temperature) * ratio


def polarizability():
    return 14.0 * A0 * A0 * A0


def build_bundle(sf, species_mass, species_natom, gas, dt, index_start=0):
    """MCCBundleElelctronInit / MCCBundleIonInit for one (species, gas) from a parsed file.

    gas: dict(mass [kg], temperature [K], density [m^-3], natom, beta_max).
    Resultant indices get gas IndexStart added (UpdateReactionIndex, MCCSigma.f90:286-299).
    """
    model = sf.model
    emin, eint, emax_exp, nsigma = grid_for_model(model)
    rtype = np.array([r.rtype for r in sf.reactions], dtype=np.int32)
    res = np.array([r.resultant + index_start for r in sf.reactions], dtype=np.int32)
    th = np.array([r.threshold for r in sf.reactions], dtype=np.float64)
    b = MCCBundle(model=model, nreaction=len(sf.reactions), nsigma=nsigma, emin=emin, einterval=eint,
                  emax=10.0 ** emax_exp, rtype=rtype, resultant=res, threshold=th,
                  gas_mass=gas["mass"], gas_temperature=gas["temperature"], beta_max=gas.get("beta_max", 0.0))
    if model in (1, 2):
        sig = sigma_normalize(sf.reactions, emin, eint, nsigma)
        b.prob, b.sigma_max = probability_nonreactive(sig, emin, eint, species_mass, gas["mass"], gas["density"])
        b.collision_ratio = 1.0 - math.exp(-b.sigma_max * dt)
    else:
        sig = sigma_normalize_reactive(th, species_natom, gas["natom"], emin, eint, nsigma)
        b.prob, b.collision_ratio = probability_reactive(sig, species_mass, gas["mass"], gas["density"],
                                                         gas.get("beta_max", 0.0), polarizability(), dt)
    return b


# ------------------------------------------------ synthetic cross sections
def log_grid(emin=0.01, emax=100.0, n=NSIGMA_RAW_MAX):
    return np.logspace(math.log10(emin), math.log10(emax), n)


def synthetic_electron_reactions(n=NSIGMA_RAW_MAX, elastic=1e-19, ion_th=16.0, ion_peak=5e-20, att_peak=5e-21,
                                 att_center=5.0, att_width=1.0, ion_product=2, att_product=1,
                                 excitation=None, dissociation=None):
    """Synthetic electron set on a log grid 0.01-100 eV: constant elastic (11),
    ionization (13) from a threshold rising to a plateau, Gaussian attachment (14)."""
    e = log_grid(n=n)
    reacs = [Reaction(0, 11, 0, 6.2293e-6, e.copy(), np.full(n, elastic))]
    if excitation is not None:
        th, peak = excitation
        s = np.where(e > th, peak * (1.0 - np.exp(-(e - th) / th)), 0.0)
        reacs.append(Reaction(0, 12, 0, th, e.copy(), s))
    ion = np.where(e > ion_th, ion_peak * (1.0 - np.exp(-(e - ion_th) / ion_th)), 0.0)
    reacs.append(Reaction(0, 13, ion_product, ion_th, e.copy(), ion))
    att = att_peak * np.exp(-0.5 * ((e - att_center) / att_width) ** 2)
    reacs.append(Reaction(0, 14, att_product, 0.0, e.copy(), att))
    if dissociation is not None:
        th, peak, code = dissociation
        s = np.where(e > th, peak * (1.0 - np.exp(-(e - th) / th)), 0.0)
        reacs.append(Reaction(0, 15, code, th, e.copy(), s))
    return reacs
