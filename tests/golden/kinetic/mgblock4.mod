: Four-state receptor with a voltage-dependent pore block, written for the tests of this
: repository: the transmitter concentration arrives through a POINTER and the block rates depend
: on the membrane voltage, so the scheme has two inputs.

NEURON {
    POINT_PROCESS MGB4
    POINTER C
    NONSPECIFIC_CURRENT i
    RANGE gmax, erev
}

UNITS {
    (nA) = (nanoamp)
    (mV) = (millivolt)
    (pS) = (picosiemens)
    (mM) = (milli/liter)
}

PARAMETER {
    gmax = 800    (pS)
    erev = 3      (mV)
    kon = 8       (/mM /ms)
    koff = 0.05   (/ms)
    beta = 0.3    (/ms)
    alpha = 0.1   (/ms)
    kb0 = 1.2     (/ms)
    ku0 = 0.4     (/ms)
    vq = 16       (mV)
}

ASSIGNED {
    v (mV)
    i (nA)
    g (pS)
    C (mM)
}

STATE { U B O OB }

BREAKPOINT {
    SOLVE kin METHOD sparse
    g = gmax * O
    i = (1e-6) * g * (v - erev)
}

FUNCTION block_on(v (mV)) (/ms) {
    block_on = kb0 * exp(-v / vq)
}

FUNCTION block_off(v (mV)) (/ms) {
    block_off = ku0 * exp(v / vq)
}

KINETIC kin {
    ~ U <-> B     (kon * C, koff)
    ~ B <-> O     (beta, alpha)
    ~ O <-> OB    (block_on(v), block_off(v))
    CONSERVE U + B + O + OB = 1
}
