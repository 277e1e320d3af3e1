"""ConstraintRecorder: the grips' displacement and reaction force, written to CSV.

Same files, columns and number formats as the reference's recorder
(constraint_recorder.py:25-138), which downstream plotting reads:

- `<scene>_constraints.csv`: step, time, load_s, then per constraint k named `name`
  the columns `c{k}_{name}_ux, _uy, _uz, _umag, _Fx, _Fy, _Fz, _Fmag`; with exactly two
  constraints two more, `u_rel_mag` (|u_1 - u_0|) and `F_axial_avg`
  ((|F_0x| + |F_1x|) / 2, the tension test's axial force);
- `<scene>_constraint_{name}.csv` per constraint (the `constraint_` prefix is not
  doubled): step, time, load_s, u_x, u_y, u_z, u_mag, F_x, F_y, F_z, F_mag.

Times and vectors are written `%.8e`, the load factor `%.6e`.  The solver supplies
`num_constraints`, `constraints_list`, `get_constraint_data()` and `load_s[None]`.
"""
import math
import os


def _e8(v):
    return "%.8e" % v


class ConstraintRecorder:
    def __init__(self, scene_name, solver, interval=1, output_dir="."):
        self.scene_name = scene_name
        self.solver = solver
        self.interval = max(1, int(interval))
        self.output_dir = output_dir
        self.num_constraints = int(getattr(solver, "num_constraints", 0))
        self.enabled = self.num_constraints > 0
        self.wide_file = None
        self.indiv_files = []
        if not self.enabled:
            return
        self.constraints_list = list(getattr(solver, "constraints_list", []))
        os.makedirs(output_dir, exist_ok=True)
        names = [self._name(k) for k in range(self.num_constraints)]
        self.wide_path = os.path.join(output_dir, "%s_constraints.csv" % scene_name)
        head = ["step", "time", "load_s"]
        for k, name in enumerate(names):
            pre = "c%d_%s" % (k, name)
            head += [pre + s for s in ("_ux", "_uy", "_uz", "_umag", "_Fx", "_Fy", "_Fz", "_Fmag")]
        if self.num_constraints == 2:
            head += ["u_rel_mag", "F_axial_avg"]
        self.wide_file = open(self.wide_path, "w", buffering=65536)
        self.wide_file.write(",".join(head) + "\n")
        self.indiv_paths = []
        for name in names:
            stem = name if name.startswith("constraint_") else "constraint_" + name
            path = os.path.join(output_dir, "%s_%s.csv" % (scene_name, stem))
            f = open(path, "w", buffering=65536)
            f.write("step,time,load_s,u_x,u_y,u_z,u_mag,F_x,F_y,F_z,F_mag\n")
            self.indiv_files.append(f)
            self.indiv_paths.append(path)
        self._last_step = -1
        print("ConstraintRecorder active: monitoring %d constraints [%s] every %d step(s) -> %s"
              % (self.num_constraints, ", ".join("'%s'" % n for n in names), self.interval,
                 self.wide_path))

    def _name(self, k):
        c = self.constraints_list[k] if k < len(self.constraints_list) else {}
        return c.get("name") or "constraint_%d" % k

    def record(self, step, time):
        """Write one row per file; a repeated step off the interval is skipped."""
        if not self.enabled:
            return
        if step % self.interval != 0 and step == self._last_step:
            return
        self._last_step = step
        data = self.solver.get_constraint_data()
        if not data:
            return
        s = float(self.solver.load_s[None])
        row = [str(step), _e8(time), "%.6e" % s]
        us, fs = [], []
        for k, d in enumerate(data):
            u = [float(x) for x in d["displacement"]]
            f = [float(x) for x in d["force"]]
            um = math.sqrt(sum(x * x for x in u))
            fm = math.sqrt(sum(x * x for x in f))
            us.append(u)
            fs.append(f)
            cols = [_e8(u[0]), _e8(u[1]), _e8(u[2]), _e8(um), _e8(f[0]), _e8(f[1]), _e8(f[2]), _e8(fm)]
            row += cols
            if k < len(self.indiv_files):
                self.indiv_files[k].write(",".join([str(step), _e8(time), "%.6e" % s] + cols) + "\n")
        if self.num_constraints == 2:
            rel = math.sqrt(sum((us[1][q] - us[0][q]) ** 2 for q in range(3)))
            row += [_e8(rel), _e8(0.5 * (abs(fs[0][0]) + abs(fs[1][0])))]
        self.wide_file.write(",".join(row) + "\n")

    def close(self):
        if not self.enabled:
            return
        if self.wide_file:
            self.wide_file.flush()
            self.wide_file.close()
            self.wide_file = None
        for f in self.indiv_files:
            f.flush()
            f.close()
        self.indiv_files = []
        print("ConstraintRecorder finished: data saved to %s" % self.wide_path)
