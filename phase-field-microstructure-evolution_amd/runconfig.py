"""Parameter plumbing around the timestep: the reference's config.yaml and
input.txt (inputs/config.yaml, tests/scaling_tests.py:53-69,
tests/run_phase_sim.py:112-128) and the parameter-sweep ensemble.

* load_config / config_list: baseline_config with each entry of a test section
  applied on top, as the reference's scaling driver builds its job list.
* write_input: input.txt in the reference's grammar, with the same fixed
  entries the reference's generator writes (ppt_rad 2 2 2, i/j/k_ppt 48 70 48,
  iseed -1, grn_rad 40, dx = dy = dz = 0.001, ndim 2).
* ensemble_members / run_ensemble: many independent simulations of uneven grid
  sizes advanced together on one GPU (pf_ensemble_step), parameters drawn from
  one ran_2 stream in member order.
"""
from __future__ import annotations

import os

import yaml


def load_config(path: str) -> dict:
    with open(path) as f:
        return yaml.safe_load(f)


def config_list(config: dict, test_type: str) -> list[dict]:
    base = config["baseline_config"]
    out = []
    for swap in config[test_type]:
        c = dict(base)
        c.update(swap)
        out.append(c)
    return out


def write_input(output_dir: str, params: dict) -> str:
    """input.txt for var_diff_hip.x (and var_diff.x) from one parameter dict."""
    p = params
    lines = [
        f"1 {p['ppt_int']}       / nrun, ppt_int ",
        f"2 2 2 {p['N_step']} {p['ifreq']}      / ppt_rad(3), N_step, ifreq",
        "48 70 48 -1        / i_ppt, j_ppt, k_ppt, iseed",
        f"40 {p['num_ppt']} {p['c_mat']} {p['c_ppt']} / grn_rad, num_ppt, c_mat, c_ppt ",
        f"{p['Nx']} {p['Ny']} {p['Nz']} {p['var']} 0.001 0.001 0.001 {p['t_step']}  / Nx, Ny, Nz, var, dx, dy, dz, t_step",
        f"{p['grad_coeff_phi']} {p['grad_coeff_c']} {p['D_S']} {p['D_L']} {p['d_gb']} {p['mob_phi']}   / grad_coeff_phi, grad_coeff_c, D_S, D_L, d_gb, mob_phi",
        f"{p['sigma_1']}  {p['sigma_2']}  / sigma_1, sigma_2",
        f"{p['am']} {p['ap']} {p['con_0_mat']} {p['con_0_ppt']} {p['gb_force']}     / am, ap, con_0_mat, con_0_ppt, gb_force",
        "2       / ndim",
        f"{p['dims'][0]} {p['dims'][1]}     / dims(2)",
        f"'{p['start']}'         / start",
    ]
    path = os.path.join(output_dir, "input.txt")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def params_from_config(pf, p: dict):
    """pf_params from one config dict (names as in input.txt / config.yaml)."""
    return pf.make_params(t_step=p["t_step"], grad_coeff_phi=p["grad_coeff_phi"],
                          grad_coeff_c=p["grad_coeff_c"], D_S=p["D_S"], D_L=p["D_L"], D_gb=p["d_gb"],
                          mob_phi=p["mob_phi"], sigma_1=p["sigma_1"], sigma_2=p["sigma_2"], am=p["am"],
                          ap=p["ap"], con_0_mat=p["con_0_mat"], con_0_ppt=p["con_0_ppt"])


# Parameter-sweep ensemble: member n has grid sizes[n % len(sizes)]^3, var
# order parameters, and kappa (both gradient coefficients) ~ U(0.25, 2),
# M = D_L ~ U(0.5, 2), L = mob_phi ~ U(0.5, 2), drawn in that order from one
# ran_2 stream in member order.
ENSEMBLE_SIZES = (64, 96, 128, 160, 192)


def ensemble_members(pf, n: int = 48, var: int = 8, sizes=ENSEMBLE_SIZES, iseed: int = -1,
                     t_step: float = 0.005):
    rng = pf.Ran2(iseed)
    out = []
    for m in range(n):
        kappa = 0.25 + 1.75 * rng()
        mob = 0.5 + 1.5 * rng()
        lmob = 0.5 + 1.5 * rng()
        g = sizes[m % len(sizes)]
        prm = pf.make_params(t_step=t_step, grad_coeff_phi=kappa, grad_coeff_c=kappa, D_L=mob, D_S=mob,
                             mob_phi=lmob)
        out.append(dict(member=m, grid=g, var=var, params=prm, kappa=kappa, M=mob, L=lmob))
    return out


def run_ensemble(pf, members, nsteps: int, mode: str = "fp32_mixed", seeds: int = 16, chunk: int = 0):
    """Create one Simulation per member (Voronoi IC, `seeds` seeds, iseed = -(m+1)),
    advance all of them nsteps together; returns the simulations (caller closes)."""
    sims = []
    for m in members:
        g = m["grid"]
        s = pf.Simulation(m["params"], g, g, g, m["var"], mode)
        s.put_voronoi(seeds, -(m["member"] + 1), 0.4, 0.02)  # = ic_voronoi + put_state, built on the device
        sims.append(s)
    step = chunk if chunk > 0 else nsteps
    done = 0
    while done < nsteps:
        k = min(step, nsteps - done)
        pf.ensemble_step(sims, k)
        done += k
    return sims
