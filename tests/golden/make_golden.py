"""Writes the golden fixtures of tests/golden/ from the reference's own CPU
path (oracle/_ref/libref.so, compiled from the reference sources by
`make -C oracle ref`). Run in the build container, where the reference is
mounted:  python tests/golden/make_golden.py [solver] [ice]
(the .npz fixtures of the solver path, the 3D-ICE exports; both by default)

Each fixture holds inputs and the reference's outputs only (arrays in .npz,
scalars in the same file). The reference itself never travels.
"""
import json
import os
import shutil
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


ICE = os.path.join(HERE, "ice")  # 3D-ICE exports of the reference's exporter


def ice_inputs():
    """Case folders for the export checks: the reference's own validation
    cases, and variants of them that reach each of the exporter's branches
    (a die with no blocks, each refusal). Written under ice/inputs/."""
    base = os.path.join(CASES, "validate_3dice")
    out = {}

    def variant(name, stack_edit=None, floor_edit=None, tsv=None):
        d = os.path.join(ICE, "inputs", name)
        os.makedirs(d, exist_ok=True)
        for f in ("stack.csv", "floorplan.csv"):
            text = open(os.path.join(base, f)).read()
            if f == "stack.csv" and stack_edit:
                text = stack_edit(text)
            if f == "floorplan.csv" and floor_edit:
                text = floor_edit(text)
            open(os.path.join(d, f), "w").write(text)
        if tsv:
            open(os.path.join(d, "tsv.csv"), "w").write(tsv)
        out[name] = d

    # tier 1 without blocks: its floorplan is the single empty element
    variant("die_without_blocks", floor_edit=lambda t: "\n".join(l for l in t.splitlines() if "tier 1" not in l) + "\n")
    variant("overlapping_blocks", floor_edit=lambda t: t + "extra, Si tier 2 (bottom), 1.5, 1.5, 2.0, 2.0, 1\n")
    variant("bottom_sink", stack_edit=lambda t: t.replace("# h_bottom, 0", "# h_bottom, 500"))
    variant("warm_ambient", stack_edit=lambda t: t.replace("# h_top, 2e4", "# h_top, 1.25e4").replace(
        "# ambient_C, 25", "# ambient_C, 40.5"))
    variant("two_plane_die", stack_edit=lambda t: t.replace("Si tier 2 (bottom), 50, 130, 1.66e6, 1, 1",
                                                            "Si tier 2 (bottom), 50, 130, 1.66e6, 2, 1"))
    variant("tsv_in_die", tsv="layer, x_mm, y_mm, w_mm, h_mm, cu_fraction\nSi tier 1 (top), 0.5, 0.5, 2.0, 2.0, 0.15\n")
    variant("tsv_stack", tsv="layer, x_mm, y_mm, w_mm, h_mm, cu_fraction\n"
                             "base, 1.03, 2.01, 3.3, 0.77, 0.4\nbond + BEOL, 0.5, 0.5, 2.0, 2.0, 0.15\n"
                             "bond + BEOL, 1.5, 1.5, 2.0, 2.0, 0.2\nCu spreader, 7.0, 7.0, 1.0, 2.5, 0.3\n")
    return out


def ice_fixtures():
    """The reference exporter's files for the cases that export, and its
    message for each it refuses (ice/exports.json lists both)."""
    if os.path.isdir(ICE):
        shutil.rmtree(ICE)
    os.makedirs(ICE)
    inputs = ice_inputs()
    runs = [("validate_tsv", os.path.join(CASES, "validate_tsv"), 100), ("validate_tsv", os.path.join(CASES, "validate_tsv"), 50),
            ("validate_3dice", os.path.join(CASES, "validate_3dice"), 100),
            ("validate_3dice", os.path.join(CASES, "validate_3dice"), 250),
            ("die_without_blocks", inputs["die_without_blocks"], 200), ("tsv_stack", inputs["tsv_stack"], 100),
            ("example", os.path.join(CASES, "example"), 100), ("tsv_under_core0", os.path.join(CASES, "tsv_under_core0"), 100),
            ("validate_3dice", os.path.join(CASES, "validate_3dice"), 300),
            ("overlapping_blocks", inputs["overlapping_blocks"], 100), ("bottom_sink", inputs["bottom_sink"], 100),
            ("warm_ambient", inputs["warm_ambient"], 125),
            ("two_plane_die", inputs["two_plane_die"], 100), ("tsv_in_die", inputs["tsv_in_die"], 100)]
    errors = {}
    for name, folder, cell in runs:
        out = os.path.join(ICE, f"{name}_{cell}um")
        os.makedirs(out)
        err = O.ref_export_3dice(folder, cell, out)
        rel = os.path.relpath(folder, HERE)
        if err is not None:
            os.rmdir(out)
            errors[f"{name}_{cell}um"] = dict(case=rel, cell_um=cell, error=err)
        else:
            errors[f"{name}_{cell}um"] = dict(case=rel, cell_um=cell, error=None,
                                              files=sorted(os.listdir(out)))
        print(f"3D-ICE export {name} at {cell} um: {err or 'ok'}")
    with open(os.path.join(ICE, "exports.json"), "w") as f:
        json.dump(errors, f, indent=1, sort_keys=True)


def main(sections):
    if not O.ref_available():
        sys.exit("oracle/_ref/libref.so missing: run `make -C oracle ref` first")
    if "ice" in sections:
        ice_fixtures()
    if "solver" not in sections:
        return

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
    # sections: "solver" (the .npz fixtures) and "ice" (3D-ICE exports); both by default
    main(sys.argv[1:] or ["solver", "ice"])
