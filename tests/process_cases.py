"""Single-column cases of CLOUDSC for the process pins (tests/test_oracle_processes*.py): an L137 column at a
fixed surface pressure with every process input zero except the ones a test sets, and an independent
restatement of the thermodynamics the expected values need (Tetens saturation over water and ice with the IFS
constants, the quadratic mixed-phase weight, the saturation specific humidity), written from the formulas'
published definitions rather than taken from the oracle."""
import numpy as np

import oracle
from openifs_amd import cloudsc as C

DT = 900.0
PS = 100000.0
KLEV = 137


def es_liq(p, t):
    """Saturation vapour pressure over water, Tetens' formula with the IFS constants (Pa)."""
    return p.r2es * np.exp(p.r3les * (t - p.rtt) / (t - p.r4les))


def es_ice(p, t):
    """Saturation vapour pressure over ice (Pa)."""
    return p.r2es * np.exp(p.r3ies * (t - p.rtt) / (t - p.r4ies))


def alpha(p, t):
    """Liquid share of the mixed phase: quadratic between RTICE and RTWAT."""
    x = (np.clip(t, p.rtice, p.rtwat) - p.rtice) / (p.rtwat - p.rtice)
    return np.minimum(1.0, x * x)


def q_of_e(p, e, pres):
    """Specific humidity at vapour pressure e (the IFS clips e/p at 0.5)."""
    x = np.minimum(e / pres, 0.5)
    return x / (1.0 - p.retv * x)


def qsat_mix(p, t, pres):
    a = alpha(p, t)
    return q_of_e(p, a * es_liq(p, t) + (1.0 - a) * es_ice(p, t), pres)


def qsat_liq(p, t, pres):
    return q_of_e(p, es_liq(p, t), pres)


def qsat_ice(p, t, pres):
    """Saturation over ice below RTT and over water above (the IFS's FOEEW split, a step at RTT)."""
    return q_of_e(p, np.where(t > p.rtt, es_liq(p, t), es_ice(p, t)), pres)


def dqdt(f, p, t, pres, h=1e-4):
    """Temperature derivative of a saturation function by a centred difference."""
    return (f(p, t + h, pres) - f(p, t - h, pres)) / (2.0 * h)


def levels():
    a_h, b_h = C.hybrid_levels(KLEV)
    paph = a_h + b_h * PS
    pap = C.standard_full_levels(a_h, b_h, PS)
    return paph, pap


def column(p, t=290.0, rh=0.5):
    """One column (KLON = 1), Fortran-ordered block arrays; isothermal T, relative humidity rh over the mixed
    phase, no cloud, no condensate, no convection, no input tendencies."""
    paph, pap = levels()
    arrs = C.allocate_block(1, KLEV, "dp")
    for n, kind, role in C.BLOCKED_ARRAYS:
        if role != "o":
            arrs[n][...] = 0
    arrs["pt"][0] = t
    arrs["pq"][0] = rh * qsat_mix(p, arrs["pt"][0], pap)
    arrs["pap"][0] = pap
    arrs["paph"][0] = paph
    return arrs


def run(p, arrs, prec="dp"):
    oracle.cloudsc_block(prec, p, 1, 1, 1, KLEV, DT, arrs)
    return arrs


def rho(p, t, pres):
    return pres / (p.rd * t)
