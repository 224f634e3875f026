: Five-state transmitter-gated receptor written for the tests of this repository: two binding
: steps, an open state and a desensitised state; the transmitter is read from outside the cell.

NEURON {
    POINT_PROCESS GLU5
    USEION glu READ gluo
    NONSPECIFIC_CURRENT i
    RANGE gmax, erev
}

UNITS {
    (nA) = (nanoamp)
    (mV) = (millivolt)
    (nS) = (nanosiemens)
    (mM) = (milli/liter)
}

PARAMETER {
    gmax = 2      (nS)
    erev = 0      (mV)
    kon = 5       (/mM /ms)
    koff = 2      (/ms)
    beta = 6      (/ms)
    alpha = 1.2   (/ms)
    kdes = 0.4    (/ms)
    krec = 0.03   (/ms)
    q10 = 2.0
}

ASSIGNED {
    v (mV)
    i (nA)
    g (nS)
    gluo (mM)
    tadj
}

STATE { R RA RA2 O D }

INITIAL {
    tadj = q10^((celsius - 22)/10)
    SOLVE scheme STEADYSTATE sparse
}

BREAKPOINT {
    SOLVE scheme METHOD sparse
    g = gmax * O
    i = (0.001) * g * (v - erev)
}

KINETIC scheme {
    ~ R <-> RA    (2 * kon * gluo * tadj, koff * tadj)
    ~ RA <-> RA2  (kon * gluo * tadj, 2 * koff * tadj)
    ~ RA2 <-> O   (beta * tadj, alpha * tadj)
    ~ RA2 <-> D   (kdes * tadj, krec * tadj)
    CONSERVE R + RA + RA2 + O + D = 1
}
