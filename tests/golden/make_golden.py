"""Writes the golden fixtures of tests/golden/ from the reference's own CPU
path (oracle/_ref/libref.so, compiled from the reference sources by
`make -C oracle ref`). Run in the build container, where the reference is
mounted:  python tests/golden/make_golden.py

Each fixture holds inputs and the reference's outputs only (arrays in .npz,
scalars in the same file). The reference itself never travels.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

CASES = os.path.join(HERE, "cases")  # copies of the reference's cases/ input folders


def save(name, **arrays):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrays)
    print(f"wrote {path} ({os.path.getsize(path) / 1024:.0f} KiB)")


def model_arrays(m):
    a = m.arrays()
    return {k: a[k] for k in ("gE", "gN", "gU", "gB", "cap", "P", "tsv_frac", "dz")}


def hotspot_ref(n, ny=None, at=(3e-3, 3e-3)):
    r = O.RefModel.default(n, ny)
    r.add_uniform_power(0, 22.0)
    r.add_hotspot(0, at[0], at[1], 1e-3, 3.0)
    r.add_uniform_power(2, 25.0)
    return r.finalize()


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libref.so missing: run `make -C oracle ref` first")

    # 1. Model arrays: hotspot stack on an odd grid; example case with TSVs;
    #    a case folder loaded by the reference's loader.
    m = hotspot_ref(13, 9)
    save("model_hotspot_13x9", nx=13, ny=9, **model_arrays(m))
    m = O.RefModel.example_case(20, 14, True)
    save("model_example_20x14", nx=20, ny=14, **model_arrays(m))
    m = O.RefModel.case(os.path.join(CASES, "validate_tsv"), 30, 25)
    save("model_case_validate_tsv_30x25", nx=30, ny=25, total_power=m.total_power(), **model_arrays(m))

    # 2. Multigrid levels of the example case (odd sizes on the way down).
    m = O.RefModel.example_case(21, 14, True)
    lv = m.levels()
    d = {"nlevels": len(lv)}
    for k, L in enumerate(lv):
        for name in ("gE", "gN", "gU", "gB", "cap"):
            d[f"L{k}_{name}"] = L[name]
        d[f"L{k}_nx"], d[f"L{k}_ny"] = L["nx"], L["ny"]
    save("levels_example_21x14", **d)

    # 3. One V-cycle (float MgReal) for a seeded residual, steady and dt = 1e-3.
    m = O.RefModel.example_case(21, 14, True)
    r = np.random.default_rng(7).standard_normal(m.n)
    save("vcycle_example_21x14", r=r, z_steady=m.precondition(r, 0.0), z_dt1e3=m.precondition(r, 1e-3))

    # 4. Steady MG-PCG: hotspot 24x24 (full field), 1e-9 and 1e-10; Jacobi.
    m = hotspot_ref(24)
    a = m.cpu_pcg(1e-10, precond="mg")
    b = m.cpu_pcg(1e-9, precond="mg")
    j = m.cpu_pcg(1e-9, precond="jacobi")
    save("steady_hotspot_24", T_mg_1e10=a["T"], iters_mg_1e10=a["iters"], relres_mg_1e10=a["relres"],
         T_mg_1e9=b["T"], iters_mg_1e9=b["iters"], T_jacobi_1e9=j["T"], iters_jacobi_1e9=j["iters"],
         true_relres_mg_1e10=m.true_relres(a["T"]), heat_out_mg_1e10=m.heat_out(a["T"]),
         total_power=m.total_power())

    # 5. Config 0 of the benchmark: 128x128x10, hotspot centred at (5, 5) mm,
    #    to 1e-10 (summary: per-plane min/max/mean and 4096 seeded cells), and
    #    the uniform-power run with its exact 1D chain.
    m = hotspot_ref(128, at=(5e-3, 5e-3))
    a = m.cpu_pcg(1e-10, precond="mg")
    T = a["T"].reshape(m.nz, -1)
    idx = np.sort(np.random.default_rng(11).choice(m.n, 4096, replace=False))
    save("steady_config0_128", iters=a["iters"], relres=a["relres"], plane_min=T.min(1), plane_max=T.max(1),
         plane_mean=T.mean(1), sample_idx=idx, sample_T=a["T"][idx], true_relres=m.true_relres(a["T"]),
         total_power=m.total_power(), solve_ms=a["solve_ms"])
    u = O.RefModel.default(128)
    u.add_uniform_power(0, 25.0)
    u.add_uniform_power(2, 25.0)
    u.finalize()
    a = u.cpu_pcg(1e-10, precond="mg")
    T = a["T"].reshape(u.nz, -1)
    save("steady_uniform_128", iters=a["iters"], plane_min=T.min(1), plane_max=T.max(1),
         analytic=u.analytic_1d())

    # 6. Transient (MG): hotspot 24x24, 20 ms with 20% growth; trajectory + final field.
    m = hotspot_ref(24)
    tr = m.run_transient("mg", t_end=0.02, growth=1.2, tol=1e-9)
    save("transient_hotspot_24", t=tr["t"], dt=tr["dt"], iters=tr["iters"], T_max=tr["T_max"], q_out=tr["q_out"],
         q_stored=tr["q_stored"], T=tr["T"], total_power=m.total_power())


if __name__ == "__main__":
    main()
