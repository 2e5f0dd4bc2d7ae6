"""Algorithmic cost constants used for roofline reporting.

FLOPs count +, -, *, /, sqrt and compares as 1 and a fused multiply-add as
2, for DIM=3 with the Wendland C4 kernel, over useful work only (pairs
inside the cut, cells actually tested).  Bytes are the compulsory FP64
structure-of-arrays traffic of a stage that reads each input and writes
each output once.
"""

# Barnes-Hut gravity walk (BHNode::calc_force, src/bhtree.cpp:301-331)
GRAV_FLOP_PP = 71     # softened particle-particle interaction (two f, two g)
GRAV_FLOP_PN = 23     # monopole of an accepted cell
GRAV_FLOP_OPEN = 10   # one opening test (every tested cell)

# Pair sums (useful pairs: r < h_i for gather, r < max(h_i, h_j) for force)
NEWTON_FLOP_PAIR = 33
DENSITY_FLOP_PAIR = 53    # DISPH final sums incl. signal velocity
BALSARA_FLOP_PAIR = 48
DISPH_FORCE_FLOP_PAIR = 115
SSPH_FORCE_FLOP_PAIR = 105
GSPH2_FORCE_FLOP_PAIR = 255

# Compulsory FP64 bytes per particle-step by stage
BYTES_PER_PARTICLE = {
    "timestep": 32, "predict": 184, "pre_interaction": 140, "set_kernel": 8,
    "fluid_force": 152, "gravity": 96, "correct": 104,
}

# MI355X published peaks
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0

# Neighbour search: each particle's FP64 position and search radius read once
# (32 B), and its neighbour set written once as 4-byte indices (one per
# symmetric pair, the set the force pass consumes)
SEARCH_BYTES_PER_PARTICLE = 32
SEARCH_BYTES_PER_PAIR = 4


def search_bytes(n: int, stats: dict, steps: int) -> float:
    """Algorithmic bytes of one neighbour search."""
    return SEARCH_BYTES_PER_PARTICLE * n + SEARCH_BYTES_PER_PAIR * stats["force_pairs"] / steps


def pair_flops(stats: dict, n: int, steps: int, sph_type: str = "disph") -> tuple[float, float]:
    """Algorithmic FLOPs per step of the pre-interaction pair pass and of the
    force pass, from the pair counters: each Newton iteration of a particle
    sweeps its gather pairs (counted as the mean number of iterations per
    particle times the gather pairs), then one sweep gives the density and
    Balsara sums; the force pass evaluates every force pair once."""
    gp = stats["gather_pairs"] / steps
    iters = stats["newton_iterations"] / steps / max(n, 1)
    pre = NEWTON_FLOP_PAIR * gp * iters + (DENSITY_FLOP_PAIR + BALSARA_FLOP_PAIR) * gp
    per_force = {"disph": DISPH_FORCE_FLOP_PAIR, "ssph": SSPH_FORCE_FLOP_PAIR, "gsph2": GSPH2_FORCE_FLOP_PAIR}[sph_type]
    return pre, per_force * stats["force_pairs"] / steps


def gravity_flops(stats: dict) -> float:
    """Algorithmic FLOPs of the gravity stage from its interaction counters."""
    tests = stats["gravity_pn"] + stats["gravity_opened"]
    return (GRAV_FLOP_PP * stats["gravity_pp"] + GRAV_FLOP_PN * stats["gravity_pn"]
            + GRAV_FLOP_OPEN * tests)
