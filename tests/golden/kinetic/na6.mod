TITLE six-state sodium channel for the kinetic front end tests

COMMENT
Written for these tests: three closed states, one open state, fast and slow inactivation.
Exercises FUNCTION, PROCEDURE with UNITSOFF, a temperature factor set in INITIAL through a
top-level LOCAL, and unit annotations inside expressions.
ENDCOMMENT

NEURON {
    SUFFIX na6
    USEION na READ ena WRITE ina
    RANGE gbar, g
}

UNITS {
    (mA) = (milliamp)
    (mV) = (millivolt)
}

PARAMETER {
    gbar = 0.08   (mho/cm2)
    a0   = 12     (/ms)
    va   = -30    (mV)
    sa   = 6      (mV)
    b0   = 0.5    (/ms)
    vb   = -50    (mV)
    sb   = 20     (mV)
    kf   = 1.2    (/ms)   : open -> fast inactivated
    kr   = 0.4    (/ms)
    vh   = -60    (mV)
    sh   = 5      (mV)
    ks   = 0.002  (/ms)
    ksr  = 0.0005 (/ms)
}

ASSIGNED {
    v    (mV)
    ena  (mV)
    ina  (mA/cm2)
    g    (mho/cm2)
    am   (/ms)
    bm   (/ms)
    hr   (/ms)
    qt
}

STATE { C3 C2 C1 O IF IS }

LOCAL tfac

INITIAL {
    qt = 2.3^((celsius - 22 (degC))/10 (degC))
    tfac = qt
    SOLVE states STEADYSTATE sparse
}

BREAKPOINT {
    SOLVE states METHOD sparse
    g = gbar * O
    ina = g * (v - ena)
}

FUNCTION expf(x, s) {
    expf = exp(x / s)
}

PROCEDURE rates(v (mV)) {
    UNITSOFF
    am = tfac * a0 / (1 + expf(-(v - va), sa))
    bm = tfac * b0 * expf(-(v - vb), sb)
    hr = tfac * kr / (1 + expf(v - vh, sh))
    UNITSON
}

KINETIC states {
    rates(v)
    ~ C3 <-> C2 (3 * am, bm)
    ~ C2 <-> C1 (2 * am, 2 * bm)
    ~ C1 <-> O  (am, 3 * bm)
    ~ O  <-> IF (tfac * kf, hr)
    ~ C1 <-> IF (0.1 * tfac * kf, hr)
    ~ IF <-> IS (tfac * ks, tfac * ksr)
    CONSERVE C3 + C2 + C1 + O + IF + IS = 1
}
