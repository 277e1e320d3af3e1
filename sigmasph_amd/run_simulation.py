"""Headless driver of a scene on the GPU step: the reference's run_simulation.py loop
(run_simulation.py:618-1337) without the GGUI window.

    python -m sigmasph_amd.run_simulation --scene_file deck.json --no-gui

What it keeps from the reference, in the same order: the deck read through the
registry, `ParticleSystem` / `build_solver()` / `initialize()`, the render-cycle loop
of `Time.stepsPerRenderUpdate` solver steps that sets `solver.dt[None]` before each
step (clipping the last step to `Time.duration`) and follows the adaptive dt after it,
`Time.maxSteps` as a cap on render cycles, the `ConstraintRecorder` CSVs, the
extended-XYZ dumps of `Output.txt` and the AMBER-NetCDF trajectory `<scene>.nc` of
`Output.netcdf` every `Output.interval` cycles from cycle 0 (one dump feeding both), the
status line once per cycle (KE / IE / SE / HE / PE / E, the path terms VW / WW / LW
with the residual R, WD, NE, the drift against E0, the off-grid count), the NaN abort,
and the closing wall-clock line, energy table, momentum table and budget terms.

What it refuses, with a message: the window (run with --no-gui), `Output.frame`
(screenshots of that window), PLY / OBJ output and `--hg-probe`, none of which this
build provides.

`main()` returns a dict (steps, render cycles, simulated time, initial and final
budgets and momenta, the stop reason) so that a test can read what the run printed.
"""
import argparse
import os
import sys
import time

import numpy as np

from .config import SimConfig
from .constraint_recorder import ConstraintRecorder
from .particle_system import ParticleSystem


def _hms(seconds):
    s = int(seconds)
    return "%d:%02d:%02d" % (s // 3600, (s // 60) % 60, s % 60)


BAR = "=" * 96
SUB = "-" * 96


def _clean(v):
    return 0.0 if abs(v) < 1e-12 else v


def energy_rows(ps, b):
    """(label, value) rows of the budget table and its total, in the reference's
    partition (format_energy_summary, run_simulation.py:126-431): a solid carries IE
    split into the stress pair-force work and the viscous work (the plastic work a
    share of the former), HE, and the integrator leftover LW as a reservoir of its
    own; a fluid carries the compression energy SE and VW; both carry KE, PE, the
    wall dissipation WW and the diffusion work WD."""
    solid = bool(ps.is_solid)
    vw = -b["visc"]
    wd = b.get("wd", b.get("gap_integrated", 0.0))
    rows = [("Kinetic Energy (KE)", b["ke"], 0), ("Potential Energy (PE)", b["pe"], 0)]
    if solid:
        rows.append(("Internal Energy (IE)", b["ie"], 0))
        rows.append(("Stress Pair-Force Work", b["ie"] - vw, 1))
        if "pw" in b:
            rows.append(("Plastic Work (PW)", b["pw"], 2))
        rows.append(("Viscous Work (VW)", vw, 1))
        rows.append(("Hourglass Work (HE)", b["hg"], 0))
        rows.append(("Integrator Leftover (LW)", -b["leftover"], 0))
        total = b["ke"] + b["pe"] + b["ie"] + b["hg"] - b["leftover"] + b["wall"]
    else:
        rows.append(("Compression Energy (SE)", b["se"], 0))
        rows.append(("Viscous Work (VW)", vw, 0))
        total = b["ke"] + b["pe"] + b["se"] + vw + b["hg"] + b["wall"] + wd
    rows.append(("Boundary Wall Dissipation (WW)", b["wall"], 0))
    rows.append(("Diffusion Work (WD)", wd, 0))
    return rows, total


def format_energy_summary(ps, b0, b1=None, step=0, t=0.0, wall=None):
    rows0, tot0 = energy_rows(ps, b0)
    e0 = abs(tot0) if abs(tot0) > 1e-12 else None
    out = [BAR]
    if b1 is None:
        out.append("ENERGY SUMMARY: INITIAL STATE (t = %.6f s, step %d)" % (t, step))
        out.append(BAR)
        out.append("  %-34s %15s   %9s" % ("Component", "Initial (t=0)", "% of |E0|"))
        out.append("  " + SUB)
        for (name, v, sub) in rows0:
            v = _clean(v)
            pct = ("%8.3f%%" % (100.0 * v / e0)) if e0 else "     --  "
            out.append("%-36s %13.4e J   %9s" % (("  " + "    " * sub + ("-> " if sub else "")) + name, v, pct))
        out.append("  " + SUB)
        out.append("  %-34s %13.4e J" % ("Total Initial Energy (E0)", tot0))
        out.append(BAR)
        return "\n".join(out)
    rows1, tot1 = energy_rows(ps, b1)
    out.append("ENERGY SUMMARY: FINAL STATE (t = %.6f s, step %d%s)"
               % (t, step, (", wall clock %s" % _hms(wall)) if wall is not None else ""))
    out.append(BAR)
    out.append("  %-34s %15s  %15s  %15s  %9s" % ("Component", "Initial (t=0)", "Final (t_end)",
                                                 "Change", "% of |E0|"))
    out.append("  " + SUB)
    for (name, v0, sub), (_, v1, _) in zip(rows0, rows1):
        d = _clean(v1 - v0)
        pct = ("%+8.3f%%" % (100.0 * d / e0)) if e0 else "     --  "
        out.append("%-36s %13.4e J  %13.4e J  %+13.4e J  %9s"
                   % (("  " + "    " * sub + ("-> " if sub else "")) + name, _clean(v0), _clean(v1), d, pct))
    out.append("  " + SUB)
    d = tot1 - tot0
    pct = ("%+8.3f%%" % (100.0 * d / e0)) if e0 else "     --  "
    out.append("%-36s %13.4e J  %13.4e J  %+13.4e J  %9s" % ("  Total Energy (E)", tot0, tot1, d, pct))
    out.append("  The integrator leftover LW is %+.4e J (first order in dt); the rest of the change is"
               % b1["leftover"])
    out.append("  the spatial residual: density diffusion, the shift and its ALE transports, the wall clamp.")
    out.append(BAR)
    return "\n".join(out)


def format_momentum_summary(ps, solver, p0, g0, p1, g1, step=0, t=0.0):
    """The momentum table (format_momentum_summary, run_simulation.py:433-600): each
    component's initial, final and change, against |p0| where that is not zero and
    against the gross momentum Sum m|v| always; in axisymmetry only p_x is conserved."""
    axi = bool(getattr(ps, "axisymmetric", False))
    scale = max(g0, g1)
    floor = 1e-9 * scale
    names = (["p_x  (axial, CONSERVED)", "p_r  (radial flux, not conserved)", "p_theta (hoop, no velocity)"]
             if axi else ["p_x", "p_y", "p_z"])
    out = [BAR, "MOMENTUM BALANCE (t = %.6f s, step %d)" % (t, step), BAR,
           "  %-34s %15s  %15s  %15s  %9s %11s" % ("Component", "Initial (t=0)", "Final (t_end)",
                                                  "Change", "% of |p0|", "% of gross"),
           "  " + SUB]

    def row(name, a, b):
        d = b - a
        pa = ("%+8.3f%%" % (100.0 * d / abs(a))) if abs(a) > floor else "     --  "
        pg = ("%+10.5f%%" % (100.0 * d / scale)) if scale > 0.0 else "       --  "
        return "  %-34s %15.4e  %15.4e  %+15.4e  %9s %11s" % (name, a, b, d, pa, pg)

    for k, nm in enumerate(names):
        out.append(row(nm, float(p0[k]), float(p1[k])))
    out.append("  " + SUB)
    if not axi:
        out.append(row("|p|", float(np.linalg.norm(p0)), float(np.linalg.norm(p1))))
    out.append("  %-34s %15.4e  %15.4e  %+15.4e" % ("gross momentum (Sum m |v|)", g0, g1, g1 - g0))
    if scale > 0.0:
        idx = [0] if axi else [0, 1, 2]
        worst = max(abs(float(p1[k]) - float(p0[k])) for k in idx)
        out.append("  The %s is %.3e, %.5f%% of the gross momentum."
                   % ("axial drift" if axi else "largest component drift", worst, 100.0 * worst / scale))
    agents = []
    if any(abs(float(c)) > 0.0 for c in solver.g):
        agents.append("gravity")
    if getattr(ps, "walls", False):
        agents.append("reflecting walls")
    if getattr(solver, "num_constraints", 0):
        agents.append("%d displacement grip(s)" % solver.num_constraints)
    if agents:
        out.append("  External agents act on this deck (%s): the total is not expected to be constant."
                   % ", ".join(agents))
    out.append(BAR)
    return "\n".join(out)


def write_xyz(path, ps, data, frame):
    """Extended-XYZ dump (run_simulation.py:1110-1137): pos, object_id, div, lam,
    pressure, and for a solid the stresses, von Mises, eps_plastic and eroded."""
    cols = [data["position"], data["object_id"], data["divergence"], data["lam"], data["pressure"]]
    props = "pos:R:3:object_id:I:1:div:R:1:lam:R:1:pressure:R:1"
    if ps.is_solid:
        cols += [data["sigma_xx"], data["sigma_yy"], data["sigma_zz"], data["sigma_xy"],
                 data["von_mises"], data["eps_plastic"], data["eroded"]]
        props += (":sigma_xx:R:1:sigma_yy:R:1:sigma_zz:R:1:sigma_xy:R:1"
                  ":von_mises:R:1:eps_plastic:R:1:eroded:R:1")
    arr = np.column_stack(cols)
    header = ("%d\nLattice=5.44 0.0 0.0 0.0 5.44 0.0 0.0 0.0 5.44 Properties=%s Time=%d"
              % (arr.shape[0], props, frame))
    fmt = ["%.3e"] * arr.shape[1]
    fmt[3] = "%d"
    np.savetxt(path, arr, header=header, comments="", fmt=fmt)


def main(argv=None, output_dir=None, quiet=False):
    ap = argparse.ArgumentParser(description="sigmaSPH scene on the MI355X step (headless)")
    ap.add_argument("--scene_file", default="", help="scene file (JSON or TOML)")
    ap.add_argument("--no-gui", dest="no_gui", action="store_true",
                    help="run without a window (the only mode of this build)")
    ap.add_argument("--hg-probe", dest="hg_probe", type=float, default=None)
    ap.add_argument("--precision", default="fp64", choices=("fp64", "mixed"))
    args = ap.parse_args(argv)
    log = (lambda *a, **k: None) if quiet else print
    if not args.no_gui:
        sys.exit("This build has no GGUI window: run with --no-gui.")
    if args.hg_probe is not None:
        sys.exit("--hg-probe needs measure_hourglass, which this build does not provide.")
    scene_path = args.scene_file
    config = SimConfig(scene_path)
    scene_name = os.path.splitext(os.path.basename(scene_path))[0] or "scene"
    odir = output_dir or "."
    c = config.get_cfg
    for key, what in (("exportFrame", "Output.frame writes PNG screenshots of the window and cannot "
                                      "be used with --no-gui"),
                      ("exportPly", "Output.ply is not part of this build"),
                      ("exportObj", "Output.obj is not part of this build")):
        if c(key):
            sys.exit(what + ".")
    substeps = int(c("numberOfStepsPerRenderUpdate") or 1)
    output_txt = bool(c("exportTxt"))
    output_netcdf = bool(c("exportNetcdf"))
    nc_writer = None
    output_interval = int(c("exportInterval") or 1)
    sim_duration = c("sim_duration")
    max_steps = c("maxSteps")

    ps = ParticleSystem(config, precision=args.precision)
    solver = ps.build_solver()
    solver.initialize()

    out_c = c("exportConstraints")
    has_cons = getattr(solver, "num_constraints", 0) > 0
    recorder = None
    if has_cons and (out_c is True or out_c is None):
        recorder = ConstraintRecorder(scene_name, solver, interval=c("exportConstraintsInterval") or 1,
                                      output_dir=odir)

    dt_adaptive = float(solver.dt[None])
    dt_cfl = float(solver.CFL * ps.h_min / solver.c_signal)
    log("dt %g   (acoustic CFL limit %.4g, dt/dt_CFL = %.4f)" % (dt_adaptive, dt_cfl, dt_adaptive / dt_cfl))
    is_fluid = not ps.is_solid
    g_on = any(abs(float(v)) > 0.0 for v in solver.g)
    b0 = solver.energy_budget()
    p0 = solver.compute_momentum()
    g0 = solver.compute_gross_momentum()
    E0 = b0["total"] if not is_fluid else b0["ke"] + b0["se"] + b0["pe"]
    has_budget = is_fluid or g_on or bool(getattr(solver, "solid_visc_work_on", False)) or bool(ps.walls)
    has_gap = (not is_fluid) and bool(getattr(solver, "compression_energy_complete", False))
    total_mass = solver.compute_total_mass()

    def off_grid():
        n = int(ps.off_grid_num[None])
        frac = float(ps.off_grid_mass[None]) / total_mass if total_mass > 0 else 0.0
        return n, frac, (solver.compute_kinetic_energy_off_grid() if n > 0 else 0.0)

    log(format_energy_summary(ps, b0, step=0, t=0.0))
    sim_time = 0.0
    cnt = 0
    total_steps = 0
    wall_t0 = time.perf_counter()
    stop = "left the loop with no stop condition"
    status = []
    while True:
        if sim_duration is not None:
            left = sim_duration - sim_time
            if left <= 0 or left < 0.1 * dt_adaptive:
                stop = "reached duration"
                log("\nSimulation finished at t = %.6f s" % sim_time)
                break
        if (output_txt or output_netcdf) and cnt % output_interval == 0:
            obj_data = ps.dump()
            if output_txt:
                write_xyz(os.path.join(odir, "dump_%04d.xyz" % cnt), ps, obj_data, cnt)
            if output_netcdf:
                # created on the first frame, then appended to and flushed every interval
                if nc_writer is None:
                    from .netcdf_writer import NetCDFTrajectoryWriter
                    nc_writer = NetCDFTrajectoryWriter(os.path.join(odir, scene_name + ".nc"), ps.domain_size,
                                                       ps.is_solid, ps.two_d)
                nc_writer.write(obj_data, sim_time)
        for _ in range(substeps):
            if sim_duration is not None:
                left = sim_duration - sim_time
                if left <= 0 or left < 0.1 * dt_adaptive:
                    break
                solver.dt[None] = left if left < dt_adaptive else dt_adaptive
            else:
                solver.dt[None] = dt_adaptive
            dt_stepped = float(solver.dt[None])
            solver.step()
            new_dt = float(solver.dt[None])
            if abs(new_dt - dt_adaptive) > 1e-10:
                if new_dt < 0.9 * dt_adaptive:
                    log("  [dt ADAPTIVE] reduced from %.2e to %.2e at t=%.4fs" % (dt_adaptive, new_dt, sim_time))
                dt_adaptive = new_dt
            sim_time += dt_stepped
            total_steps += 1
            if recorder:
                recorder.record(total_steps, sim_time)
        solver.dt[None] = dt_adaptive
        # the status line (print_status, run_simulation.py:843-932)
        b = solver.energy_budget()
        ke, ie, se, he, pe = b["ke"], b["ie"], b["se"], b["hg"], b["pe"]
        tot = (ke + se + pe) if is_fluid else (ke + ie + he + pe)
        if not np.all(np.isfinite([ke, ie, se, he, pe, tot])):
            log("\n[FATAL] NaN detected in energy at t=%.4fs (step %d): KE=%s, IE=%s, SE=%s, HE=%s, PE=%s"
                % (sim_time, cnt, ke, ie, se, he, pe))
            log("Simulation aborted due to numerical instability (NaN in energy).")
            sys.exit(1)
        de = (("dE: %+.3f%%" % (100.0 * (tot - E0) / abs(E0))) if abs(E0) > 1e-12
              else "dE: %+.3e" % (tot - E0))
        s = "KE: %.4e" % ke
        if not is_fluid:
            s += " | IE: %.4e" % ie
        s += " | SE: %.4e" % se
        if not is_fluid:
            s += " | HE: %.4e" % he
        if is_fluid or g_on:
            s += " | PE: %.4e" % pe
        s += " | E: %.4e" % tot
        if has_budget:
            vw, ww, lw = b["visc"], b["wall"], b["leftover"]
            resid = ((tot - E0) - vw + ww - lw) if is_fluid else ((tot - E0) + ww - lw)
            s += " | VW: %.3e | WW: %.3e | LW: %+.3e | R: %+.3e" % (vw, ww, lw, resid)
            wd = b.get("wd", 0.0)
            if abs(wd) > 1e-12:
                s += " | WD: %+.3e" % wd
            elif has_gap:
                s += " | WD: %+.3e" % (ie + vw - b.get("comp", 0.0))
            if getattr(solver, "e_guard_on", False):
                s += " | NE: %+.3e" % b.get("e_neg", 0.0)
        n_off, frac_off, ke_off = off_grid()
        off = (" | off-grid: %d (%.2f%% mass, KE %.3e)" % (n_off, 100.0 * frac_off, ke_off)) if n_off else ""
        line = ("[t=%.4fs] [Step %d] [wall %s] dt/dt_CFL: %.4f | %s | %s%s"
                % (sim_time, cnt, _hms(time.perf_counter() - wall_t0), dt_adaptive / dt_cfl, s, de, off))
        status.append(line)
        log(line)
        cnt += 1
        if max_steps is not None and cnt >= max_steps:
            stop = "hit maxSteps"
            break
    if recorder:
        recorder.close()
    if nc_writer is not None:
        nc_writer.close()
        log("[NETCDF] %d frames written to %s.nc" % (nc_writer.frames_written, scene_name))
    wall = time.perf_counter() - wall_t0
    rate = total_steps / wall if wall > 0 else 0.0
    log("\n[WALL CLOCK] %s (%.2f s) for %d solver steps over %d render cycles -- %.0f steps/s"
        % (_hms(wall), wall, total_steps, cnt, rate))
    log("             stopped: %s, at t = %.6g s%s"
        % (stop, sim_time, (" of %g" % sim_duration) if sim_duration is not None else ""))
    b1 = solver.energy_budget()
    p1 = solver.compute_momentum()
    g1 = solver.compute_gross_momentum()
    log(format_energy_summary(ps, b0, b1, step=total_steps, t=sim_time, wall=wall))
    log(format_momentum_summary(ps, solver, p0, g0, p1, g1, step=total_steps, t=sim_time))
    if not is_fluid:
        e0 = abs(b0["total"]) or 1.0
        log("\nENERGY BUDGET TERMS (change since t = 0, and as %% of |E0| = %.4e J):" % e0)
        for k, v in b1.items():
            if isinstance(v, bool) or not isinstance(v, (int, float)):
                continue
            d = float(v) - float(b0.get(k, 0.0))
            log("  %-16s %+.4e J   %+8.3f%%" % (k, d, 100.0 * d / e0))
    result = {"steps": total_steps, "cycles": cnt, "sim_time": sim_time, "stop": stop,
              "budget0": b0, "budget1": b1, "p0": p0, "p1": p1, "gross0": g0, "gross1": g1,
              "status": status, "wall": wall, "dt": dt_adaptive}
    ps.close()
    return result


if __name__ == "__main__":
    main()
