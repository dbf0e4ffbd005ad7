"""dpnegf_amd -- MI355X-native (gfx950) NEGF hot path: recursive Green's
function, Lopez-Sancho lead self-energies, transmission and Ozaki density,
as HIP kernels behind a C ABI, with the ``dpnegf.negf`` Python interface."""
__version__ = "0.1.0"
