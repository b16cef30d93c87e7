"""Golden cases of the measurement layer (observables.hpp:36-101), shared by
the oracle and GPU tests. The reference harness measured them on a
thermalised pure-gauge field of each anisotropic lattice; the fixture keys are
<tag>_W_<label>_<Ra>x<Rb>, <tag>_Wfund_<Ra>x<Rb> (wilson_loop_fund),
<tag>_P_<label>_t<dir> and <tag>_rect (wilson_rect at x=5, mu=0, nu=D-1, 2x3)."""

# probe label in the fixture keys -> rep spec
SPEC = {"fund": "fund", "adj": "adj", "2": "2", "11": "1,1", "21": "2,1", "3": "3", "22": "2,2", "61": "6,1",
        "4": "4"}
SIZES = [(1, 1), (1, 2), (2, 1), (2, 2), (1, 3), (3, 2)]
# (tag, D, N, extents, probe labels)
CASES = [
    ("loops_su3", 4, 3, (4, 3, 4, 5), ["fund", "adj", "2", "11", "21", "3", "22", "61"]),
    ("loops_su2", 3, 2, (5, 4, 3), ["fund", "adj", "2", "3", "4"]),
]


def sizes_for(label):
    """The {6,1} probe (d = 48) was measured on loops of area <= 4 only."""
    return [s for s in SIZES if label != "61" or s[0] * s[1] <= 4]
