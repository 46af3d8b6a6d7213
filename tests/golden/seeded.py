"""Seeded initial temperature for the variable-viscosity workloads (numpy).

T_i = (1 - z_i) + 0.01 * r_i * sin(pi * z_i), r_i = 2 u_i - 1, where u_i is the
top 53 bits of SplitMix64 applied once to the state ``seed + node_id``, scaled by
2^-53.  node_id = iz + ix * (nz + 1) is the reference's global node number
(Construct_ien_id_lm.c:28-37), so the array is in the reference's sorted-node
order, the order read by Initial_condition.c:171-195.  Bottom and top rows are
then forced to 1 and 0 (Initial_condition.c:122-135 does the same for 'T' nodes).
Height is 1 (unit box or aspect-ratio box of height 1).
"""
import numpy as np

SEED = 20261202
_M = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(state):
    """One SplitMix64 output for each uint64 state (Steele, Lea, Flood 2014)."""
    with np.errstate(over="ignore"):
        z = (np.asarray(state, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15)) & _M
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M
        return z ^ (z >> np.uint64(31))


def seeded_temperature(nx, nz, seed=SEED, amp=0.01):
    """Nodal T for an nx x nz element box, array index = node id."""
    gnoz = nz + 1
    ix, iz = np.meshgrid(np.arange(nx + 1), np.arange(nz + 1), indexing="ij")
    node = (iz + ix * gnoz).ravel().astype(np.uint64)
    z = (iz.ravel() * (1.0 / nz)).astype(np.float64)
    u = (splitmix64(np.uint64(seed) + node) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    r = 2.0 * u - 1.0
    T = (1.0 - z) + amp * r * np.sin(np.pi * z)
    T[iz.ravel() == 0] = 1.0
    T[iz.ravel() == nz] = 0.0
    return T
