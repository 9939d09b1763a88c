"""Synthetic stacks of the benchmark workloads, seeded and reproducible.

Every workload is a *case spec* (layers, power blocks, TSV regions, in the
units of the reference's case folders, case3d.h:1-13). A spec can be built
into a model through the C ABI, or written out as a case folder, so the
reference's own loader reads the identical inputs.

Random numbers come from splitmix64 (Steele, Lea and Flood, "Fast
Splittable Pseudorandom Number Generators", OOPSLA 2014): a 64-bit state
advanced by 0x9E3779B97F4A7C15 and mixed by two xor-shift-multiply rounds.

Choices the workload descriptions leave open:
* Background power: 10% of each die's 25 W is spread uniformly; the 12
  blocks are rescaled to the other 90% (22.5 W).
* TSV clusters: 16 x 16 cells of the 1024 x 1024 grid (0.15625 mm squares on
  a 64 x 64 cluster lattice), copper fraction 1 (k = 400, Cu rho*c), in the
  die's Si layer and the bond layer. The cluster count per die is chosen so
  the clusters cover ~5% of the die: round(0.05 * 64 * 64) = 205 (the "8
  clusters" reading would cover 0.2%). The same physical clusters are used
  at every grid size.
"""
import math
import os

from . import DEFAULT_STACK, Model

MASK64 = (1 << 64) - 1
DIE = 10e-3  # 10 mm x 10 mm footprint


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & MASK64

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / (1 << 53)))

    def below(self, n):
        return self.next() % n


def config0_layers():
    """The reference's default stack (stack3d.h:153-162), 10 z-planes."""
    return list(DEFAULT_STACK)


def config2_layers():
    """Same layers with z-cells 4/2/4/2/4: 16 z-planes."""
    planes = (4, 2, 4, 2, 4)
    return [(name, t, k, rc, p, act) for (name, t, k, rc, _, act), p in zip(DEFAULT_STACK, planes)]


def random_blocks(seed, count=12, side=(0.5e-3, 3e-3), watts=(0.5, 8.0), total=22.5, tries=100000):
    """`count` non-overlapping rectangles with sides U[side], power
    log-uniform in [watts], rescaled to `total` W. Rejection sampling."""
    rng = SplitMix64(seed)
    out = []
    n = 0
    while len(out) < count:
        n += 1
        if n > tries:
            raise RuntimeError("could not place the blocks")
        w, h = rng.uniform(*side), rng.uniform(*side)
        x0, y0 = rng.uniform(0.0, DIE - w), rng.uniform(0.0, DIE - h)
        if any(x0 < b[0] + b[2] and b[0] < x0 + w and y0 < b[1] + b[3] and b[1] < y0 + h for b in out):
            continue
        p = math.exp(rng.uniform(math.log(watts[0]), math.log(watts[1])))
        out.append([x0, y0, w, h, p])
    s = sum(b[4] for b in out)
    for b in out:
        b[4] *= total / s
    return [tuple(b) for b in out]


def tsv_clusters(seed, count, lattice=64):
    """`count` distinct cells of a lattice x lattice grid over the die, as
    squares (x0, y0, w, h)."""
    rng = SplitMix64(seed)
    side = DIE / lattice
    picked = set()
    while len(picked) < count:
        picked.add(rng.below(lattice * lattice))
    return [((q % lattice) * side, (q // lattice) * side, side, side) for q in sorted(picked)]


def config2_spec(seed=1000, tsv=True, cluster_count=205, active=(True, True), fraction_seed=None):
    """The 1024^2 x 16 TSV stack of the benchmark: per die 12 random blocks
    (seed + die) on a 10% uniform background, 25 W per die, and TSV clusters.
    `active` switches a die's power off (both on by default). fraction_seed:
    each cluster's copper fraction drawn from U[0.05, 1] instead of 1 (on a
    grid whose cells the clusters cover in part, no two columns then need be
    alike)."""
    layers = config2_layers()
    die_layer = (0, 2)  # Si tier 2 (bottom), Si tier 1 (top)
    bond = 1
    blocks, tsvs, uniform = [], [], []
    for d in (0, 1):
        if not active[d]:
            continue
        for k, b in enumerate(random_blocks(seed + d)):
            blocks.append((f"die{d}_b{k}", die_layer[d]) + b)
        uniform.append((die_layer[d], 2.5))
        if tsv:
            frac = SplitMix64(fraction_seed + d) if fraction_seed is not None else None
            for c in tsv_clusters(seed + 1000 + d, cluster_count):
                f = frac.uniform(0.05, 1.0) if frac else 1.0
                tsvs.append((die_layer[d],) + c + (f,))
                tsvs.append((bond,) + c + (f,))
    return dict(layers=layers, blocks=blocks, uniform=uniform, tsvs=tsvs, h_top=2e4, h_bot=0.0, Lx=DIE, Ly=DIE)


def config3_instances(seed=3000, count=64, sizes=(128, 256, 512), p_on=0.6):
    """The batched workload: `count` (n, spec) pairs on the 10-plane default
    stack, n drawn from `sizes` with equal probability; each of the 12
    blocks per die of the 1024^2 TSV stack's floorplan (seed 1000, with its
    TSV clusters and 2.5 W background per die) switched on with probability
    p_on. All draws from one splitmix64 stream."""
    rng = SplitMix64(seed)
    base = config2_spec(1000)
    out = []
    for _ in range(count):
        n = sizes[rng.below(len(sizes))]
        blocks = [b for b in base["blocks"] if rng.uniform() < p_on]
        spec = dict(base, layers=config0_layers(), blocks=blocks)
        out.append((n, spec))
    return out


def config0_spec(hotspot_at=(5e-3, 5e-3), uniform_only=False):
    """Config 0: the default stack, tier 2 with 22 W uniform + 3 W in a
    centred 1 mm square, tier 1 with 25 W uniform (or 25 + 25 W uniform)."""
    if uniform_only:
        return dict(layers=config0_layers(), blocks=[], uniform=[(0, 25.0), (2, 25.0)], tsvs=[], h_top=2e4,
                    h_bot=0.0, Lx=DIE, Ly=DIE, hotspots=[])
    return dict(layers=config0_layers(), blocks=[], uniform=[(0, 22.0), (2, 25.0)], tsvs=[], h_top=2e4, h_bot=0.0,
                Lx=DIE, Ly=DIE, hotspots=[(0, hotspot_at[0], hotspot_at[1], 1e-3, 3.0)])


def build(spec, n, ny=None):
    """Finalized model of a spec on an n x n (or n x ny) grid."""
    m = Model.from_layers(spec["layers"], n, ny or n, spec["Lx"], spec["Ly"], spec["h_top"], spec["h_bot"])
    for layer, watts in spec["uniform"]:
        m.add_uniform_power(layer, watts)
    for b in spec["blocks"]:
        m.add_power_block(b[1], b[2], b[3], b[4], b[5], b[6])
    for h in spec.get("hotspots", []):
        m.add_hotspot(*h)
    for t in spec["tsvs"]:
        m.add_tsv_region(*t)
    return m.finalize()


def write_case(spec, folder):
    """Writes the spec as a case folder (stack.csv, floorplan.csv, tsv.csv),
    with full double precision. Uniform power becomes a full-die block;
    hotspots (cell-centre rule) have no case-file form and are refused."""
    if spec.get("hotspots"):
        raise ValueError("hotspots cannot be written to a case folder")
    os.makedirs(folder, exist_ok=True)
    names = [l[0] for l in spec["layers"]]
    with open(os.path.join(folder, "stack.csv"), "w") as f:
        f.write(f"# die_mm, {spec['Lx'] * 1e3!r}, {spec['Ly'] * 1e3!r}\n# h_top, {spec['h_top']!r}\n")
        f.write(f"# h_bottom, {spec['h_bot']!r}\n# ambient_C, 25\n")
        f.write("name, thickness_um, k_W_per_mK, rho_c_J_per_m3K, planes, active\n")
        for name, t, k, rc, p, act in spec["layers"]:
            f.write(f"{name}, {t * 1e6!r}, {k!r}, {rc!r}, {p}, {int(act)}\n")
    with open(os.path.join(folder, "floorplan.csv"), "w") as f:
        f.write("name, layer, x_mm, y_mm, w_mm, h_mm, power_W\n")
        for layer, watts in spec["uniform"]:
            f.write(f"background{layer}, {names[layer]}, 0, 0, {spec['Lx'] * 1e3!r}, {spec['Ly'] * 1e3!r}, {watts!r}\n")
        for b in spec["blocks"]:
            f.write(f"{b[0]}, {names[b[1]]}, {b[2] * 1e3!r}, {b[3] * 1e3!r}, {b[4] * 1e3!r}, {b[5] * 1e3!r}, {b[6]!r}\n")
    with open(os.path.join(folder, "tsv.csv"), "w") as f:
        f.write("layer, x_mm, y_mm, w_mm, h_mm, cu_fraction\n")
        for t in spec["tsvs"]:
            f.write(f"{names[t[0]]}, {t[1] * 1e3!r}, {t[2] * 1e3!r}, {t[3] * 1e3!r}, {t[4] * 1e3!r}, {t[5]!r}\n")
    return folder
