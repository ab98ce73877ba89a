"""tea.in decks, including the seeded random-rectangle problems.

The grammar is TeaLeaf's (TeaLeaf/tea.in, TeaLeaf/parse_config.c).  Every
deck here uses the reference physics: a 10 x 10 domain, halo depth 2,
dt = 0.004, background state density 100 / energy 1e-4 and the four
rectangles of TeaLeaf/tea.in, optionally followed by seeded random
rectangles.

Random rectangles come from one splitmix64 stream per problem, seeded with
``20261202 + config_id*1000 + problem_id``; per rectangle the draws are, in
order, two x corners, two y corners (each pair sorted), the density and the
energy.  Uniform doubles are ``(next >> 11) * 2**-53``.
"""

import math

MASK64 = (1 << 64) - 1

TEA_IN_STATES = [
    dict(density=100.0, energy=0.0001),
    dict(density=0.1, energy=25.0, xmin=0.0, xmax=1.0, ymin=1.0, ymax=2.0),
    dict(density=0.1, energy=0.1, xmin=1.0, xmax=6.0, ymin=1.0, ymax=2.0),
    dict(density=0.1, energy=0.1, xmin=5.0, xmax=6.0, ymin=1.0, ymax=8.0),
    dict(density=0.1, energy=0.1, xmin=5.0, xmax=10.0, ymin=7.0, ymax=8.0),
]


def splitmix64(seed):
    """Generator of 64-bit splitmix64 outputs (Steele, Lea & Flood 2014)."""
    state = seed & MASK64
    while True:
        state = (state + 0x9E3779B97F4A7C15) & MASK64
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        yield z ^ (z >> 31)


def _uniform(gen):
    return (next(gen) >> 11) * (1.0 / (1 << 53))


def problem_seed(config_id, problem_id=0):
    return 20261202 + config_id * 1000 + problem_id


def seeded_rectangles(count, seed, density="uniform", density_range=(0.1, 100.0),
                      energy_range=(0.1, 25.0), domain=(0.0, 10.0)):
    """``count`` random axis-aligned rectangles as state dicts.

    density="uniform": density ~ U[density_range];
    density="log10": log10(density) ~ U[density_range].
    """
    gen = splitmix64(seed)
    lo, hi = domain
    rects = []
    for _ in range(count):
        xa, xb = sorted(lo + (hi - lo) * _uniform(gen) for _ in range(2))
        ya, yb = sorted(lo + (hi - lo) * _uniform(gen) for _ in range(2))
        u = _uniform(gen)
        a, b = density_range
        rho = a + (b - a) * u if density == "uniform" else 10.0 ** (a + (b - a) * u)
        e = energy_range[0] + (energy_range[1] - energy_range[0]) * _uniform(gen)
        rects.append(dict(density=rho, energy=e, xmin=xa, xmax=xb, ymin=ya, ymax=yb))
    return rects


def _num(v):
    r = repr(float(v))
    if r.startswith("-"):
        raise ValueError("tea.in values cannot be negative (the parser drops '-')")
    return r


class Deck:
    """A tea.in deck."""

    def __init__(self, x_cells, y_cells, end_step=10, coefficient="density",
                 eps=1e-15, max_iters=10000, states=None, dt=0.004,
                 summary_frequency=10, extra=(), comment=None, extent=10.0):
        self.x_cells, self.y_cells = x_cells, y_cells
        self.extent = extent
        self.end_step = end_step
        self.coefficient = coefficient
        self.eps, self.max_iters = eps, max_iters
        self.states = list(states) if states is not None else list(TEA_IN_STATES)
        self.dt = dt
        self.summary_frequency = summary_frequency
        self.extra = list(extra)
        self.comment = comment

    @classmethod
    def with_random_rectangles(cls, x_cells, y_cells, count, seed, density="uniform",
                               density_range=(0.1, 100.0), extent=10.0, **kw):
        """The reference states plus ``count`` seeded rectangles; with
        ``extent`` != 10 the domain is extent x extent and every rectangle
        (those of tea.in included) is scaled with it."""
        f = extent / 10.0
        scaled = [dict(s) for s in TEA_IN_STATES]
        for s in scaled[1:]:
            for k in ("xmin", "xmax", "ymin", "ymax"):
                s[k] = s[k] * f
        states = scaled + seeded_rectangles(
            count, seed, density=density, density_range=density_range,
            domain=(0.0, extent))
        kw.setdefault("comment", f"seed {seed}: {count} splitmix64 rectangles")
        return cls(x_cells, y_cells, states=states, extent=extent, **kw)

    def render(self):
        lines = ["*tea"]
        if self.comment:
            lines.append(f"! {self.comment}")
        for i, s in enumerate(self.states, start=1):
            line = f"state {i} density={_num(s['density'])} energy={_num(s['energy'])}"
            if i > 1:
                line += (f" geometry=rectangle xmin={_num(s['xmin'])} xmax={_num(s['xmax'])}"
                         f" ymin={_num(s['ymin'])} ymax={_num(s['ymax'])}")
            lines.append(line)
        lines += [
            f"x_cells={self.x_cells}", f"y_cells={self.y_cells}",
            "xmin=0.0", "ymin=0.0", f"xmax={_num(self.extent)}", f"ymax={_num(self.extent)}",
            f"initial_timestep={_num(self.dt)}",
            f"end_step={self.end_step}",
            f"max_iters={self.max_iters}",
            f"summary_frequency={self.summary_frequency}",
            "use_cg",
            f"eps {self.eps!r}",
            "coefficient_inverse_density" if self.coefficient == "inverse"
            else "coefficient_density",
            "use_c_kernels",
        ]
        lines += list(self.extra)
        lines.append("*endtea")
        return "\n".join(lines) + "\n"

    def write(self, path):
        with open(path, "w") as f:
            f.write(self.render())
        return path


def config_deck(config_id, x_cells=None, end_step=None, problem_id=0):
    """Decks of the benchmark configurations.

    1: 512^2, 10 steps, K = density, the tea.in rectangles only.
    2: 4096^2, 5 steps, K = 1/density, + 16 rectangles, density U[0.1, 100].
    """
    if config_id == 1:
        n = x_cells or 512
        return Deck(n, n, end_step=end_step or 10, coefficient="density")
    if config_id == 2:
        n = x_cells or 4096
        return Deck.with_random_rectangles(
            n, n, 16, problem_seed(2, problem_id), density="uniform",
            density_range=(0.1, 100.0), end_step=end_step or 5,
            coefficient="inverse")
    raise ValueError(f"no deck for configuration {config_id}")


def float_close(a, b, rel):
    return math.isclose(a, b, rel_tol=rel, abs_tol=0.0)


# Decks of the time-to-solution field for one solver: 64 seeded rectangles,
# log10 density in [-2, 2], K = 1/density, one timestep to 1e-10.
SOLVER_KEYS = {
    "cg": [],
    "mpcg": ["use_mpcg"],
    "jpcg": ["use_jacobi_pcg"],
    "fcg": ["use_fcg", "fcg_weights=net.bin"],
    "fcgs": ["use_fcg", "fcg_coarse_sweeps=6"],
}


def tts_deck(solver, cells=4096, max_iters=12000, seed_id=3000):
    return Deck.with_random_rectangles(
        cells, cells, 64, problem_seed(5, seed_id), density="log10",
        density_range=(-2.0, 2.0), end_step=1, coefficient="inverse",
        max_iters=max_iters, extra=["rel_eps=1e-10"] + SOLVER_KEYS[solver])
