"""Output cadence of the reference's main loop: Output::Output and Output::CheckForWrites
(src/output/output.cpp:12-160, :162-320) for the VTK snapshots and restart dumps, with
the file numbering of Vtk::Write (vtk.cpp:293-354, "<base>.%04d.vtk") and Dump::Write
(dump.cpp:826-839, "dump.%04d.dmp", the counter advanced before the write), and
Output::RestartFromDump (output.cpp:325-337).

A dump also carries the scalars the reference registers for restarts (output.cpp:148-150,
vtk.cpp:285, dump.cpp:210-212): vtkLast, dumpLast, analysisLast, vtkFileNumber,
dumpFileNumber, geometry and periodicity.
"""
import os
import time

from . import dump as dmp
from . import vtk

INT, DOUBLE = 2, 0


class Output:
    def __init__(self, output_block, g, directory=".", force_no_write=False, periodic=None):
        """output_block: the [Output] entries {name: [values]} of the input file."""
        self.dir = directory
        self.force_no_write = force_no_write
        self.periodic = periodic if periodic is not None else (True, True, True)
        ob = output_block or {}
        self.vtk_enabled = False
        self.vtk_period = 0.0
        self.vtk_last = 0.0
        if "vtk" in ob:
            self.vtk_period = float(ob["vtk"][0])
            if self.vtk_period >= 0.0:  # a negative period means no file
                self.vtk_last = g.t - self.vtk_period
                self.vtk_enabled = True
        self.dump_enabled = False
        self.dump_period = 0.0
        self.dump_last = 0.0
        self.dump_time_period = -1.0
        self.analysis_last = 0.0
        if "dmp" in ob:
            self.dump_period = float(ob["dmp"][0])
            self.dump_last = g.t - self.dump_period
            if len(ob["dmp"]) > 1:
                s = ob["dmp"][1]
                unit = {"s": 1.0, "m": 60.0, "h": 3600.0, "d": 86400.0}
                if s[-1:] not in unit:
                    raise ValueError("The dump time period unit should be either s, m, h or d")
                self.dump_time_period = float(s[:-1]) * unit[s[-1]]
                self.dump_time_last = time.monotonic()
            self.dump_enabled = not (self.dump_period < 0.0 and self.dump_time_period < 0.0)
        self.vtk_number = 0
        self.dump_number = 0
        self.files = []

    # ---- writers ------------------------------------------------------------------
    def _write_vtk(self, g):
        name = os.path.join(self.dir, "data.%04d.vtk" % self.vtk_number)
        self.vtk_number += 1
        vtk.write_vtk(name, g, periodic=self.periodic)
        self.files.append(name)

    def _write_dump(self, g):
        name = os.path.join(self.dir, "dump.%04d.dmp" % self.dump_number)
        self.dump_number += 1
        extra = {"vtkLast": (self.vtk_last, DOUBLE), "dumpLast": (self.dump_last, DOUBLE),
                 "analysisLast": (self.analysis_last, DOUBLE),
                 "vtkFileNumber": (self.vtk_number, INT),
                 "dumpFileNumber": (self.dump_number, INT),
                 "geometry": (vtk.GEOMETRY_CARTESIAN, INT),
                 "periodicity": ([1 if p else 0 for p in self.periodic], INT)}
        dmp.write_dump(name, g, extra=extra)
        self.files.append(name)

    @staticmethod
    def _catch_up(last, period, t):
        """Moves `last` past outputs that should already have happened."""
        if last + period <= t and period > 0.0:
            while last <= t - period:
                last += period
        return last

    def check_for_writes(self, g):
        """Output::CheckForWrites: returns the number of files written."""
        if self.force_no_write:
            return 0
        n = 0
        if self.vtk_enabled and g.t >= self.vtk_last + self.vtk_period:
            self.vtk_last += self.vtk_period
            self._write_vtk(g)
            n += 1
            self.vtk_last = self._catch_up(self.vtk_last, self.vtk_period, g.t)
        if self.dump_enabled:
            periodic_dump = clock_dump = False
            if self.dump_period > 0.0 and g.t >= self.dump_last + self.dump_period:
                periodic_dump = True
                self.dump_last += self.dump_period
            if self.dump_time_period > 0.0:
                now = time.monotonic()
                if now - self.dump_time_last > self.dump_time_period:
                    clock_dump = True
                    self.dump_time_last = now
            # last, since a dump records the state of the other outputs
            if periodic_dump or clock_dump:
                self._write_dump(g)
                n += 1
                self.dump_last = self._catch_up(self.dump_last, self.dump_period, g.t)
        return n

    def force_write_vtk(self, g):
        """Output::ForceWriteVtk (output.cpp:341-362): the emergency snapshot."""
        if not self.force_no_write:
            self.vtk_last += self.vtk_period
            self._write_vtk(g)

    def force_write_dump(self, g):
        if not self.force_no_write:
            self._write_dump(g)

    # ---- restart ------------------------------------------------------------------
    def restart_from_dump(self, g, number=-1):
        """Output::RestartFromDump; number -1 takes the highest-numbered dump of the
        directory (dump.cpp:598-627)."""
        if number < 0:
            nums = [int(f[5:9]) for f in os.listdir(self.dir)
                    if f.startswith("dump.") and f.endswith(".dmp") and f[5:9].isdigit()]
            if not nums:
                return False
            number = max(nums)
        name = os.path.join(self.dir, "dump.%04d.dmp" % number)
        if not os.path.exists(name):
            return False
        F = dmp.restart(g, name)
        if "vtkLast" in F:
            self.vtk_last = float(F["vtkLast"][0])
        if "dumpLast" in F:
            self.dump_last = float(F["dumpLast"][0])
        if "analysisLast" in F:
            self.analysis_last = float(F["analysisLast"][0])
        if "vtkFileNumber" in F:
            self.vtk_number = int(F["vtkFileNumber"][0])
        # the counters are registered variables: the file holds the values advanced
        # before its own write
        self.dump_number = int(F["dumpFileNumber"][0]) if "dumpFileNumber" in F else number + 1
        return True
