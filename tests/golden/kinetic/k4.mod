TITLE four-state potassium channel for the kinetic front end tests

COMMENT
Written for these tests: two closed states, open, inactivated.  Voltage dependence through an
exponent z F / (R T) computed in INITIAL; the current carries a 1e-3 unit factor, so gbar in
mS/cm2 gives S/cm2.
ENDCOMMENT

NEURON {
    SUFFIX k4
    USEION k READ ek WRITE ik
    RANGE gbar
}

UNITS { (mV) = (millivolt) }

PARAMETER {
    gbar  = 24     (millimho/cm2)
    k0    = 0.05   (/ms)
    z     = 2.1    ()
    vhalf = -25    (mV)
    kin   = 0.0015 (/ms)
    kout  = 0.0002 (/ms)
    Hk    = 0.06   (/degC)
}

ASSIGNED {
    v  (mV)
    ek (mV)
    ik (milliamp/cm2)
    g  (millimho/cm2)
    kf (/ms)
    kb (/ms)
}

STATE { C2 C1 O I }

LOCAL F, RT, tf

INITIAL {
    F  = 96.485 (joule/mV)
    RT = 8.314 (joule/degC) * (celsius + 273.15)
    tf = exp(Hk * (celsius - 25))
    SOLVE kin STEADYSTATE sparse
}

BREAKPOINT {
    SOLVE kin METHOD sparse
    g  = gbar * O
    ik = g * (v - ek) * (1e-3)
}

PROCEDURE rates(v (millivolt)) {
    kf = tf * k0 * exp(0.5 * z * F / RT * (v - vhalf))
    kb = tf * k0 * exp(-0.5 * z * F / RT * (v - vhalf))
}

KINETIC kin {
    rates(v)
    ~ C2 <-> C1 (2 * kf, kb)
    ~ C1 <-> O  (kf, 2 * kb)
    ~ O  <-> I  (tf * kin, tf * kout)
    CONSERVE C2 + C1 + O + I = 1
}
