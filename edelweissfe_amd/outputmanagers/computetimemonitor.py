"""Compute time per increment (computetimemonitor.py:50-132): the difference of the
solver's running computationTimes since the previous increment, per timing category and
in total.  Printed through the journal when there is one and, with ``export=<file>``,
appended to a file: a header per step, then one line per increment with the increment
number, the time at its end, its total compute time and one column per category.

On the device the categories time the host side of each phase: a phase that ends with a
readback (the element pass's cutback flag, the convergence norms, the Krylov status)
includes its GPU work; queued launches are charged to the phase that waits for them."""

from collections import defaultdict

from .base import OutputManagerBase
from .timing import createTimingDict, timingTypes


class OutputManager(OutputManagerBase):
    identification = "ComputeTimeMonitor"

    def __init__(self, name=None, model=None, fieldOutputController=None, journal=None, plotter=None):
        self.journal = journal
        self.computingTimesOld = defaultdict(float)
        self.stepcounter = 0
        self.exportFile = None
        self.increments = []   # (increment, time end, total, {category: seconds}) per finalized increment

    def updateDefinition(self, **kwargs):
        self.exportFile = kwargs.get("export")
        if self.exportFile is not None:
            with open(self.exportFile, "w") as f:
                f.write("# \n# EdelweissFE: computing times per increment\n#\n")

    def initializeStep(self, step):
        self.computingTimesOld = createTimingDict()
        self.stepcounter += 1
        if self.exportFile is not None:
            with open(self.exportFile, "a") as f:
                f.write("#\n# simulation step {:}\n#\n#".format(self.stepcounter))
                f.write(" {:<20} {:<20} {:<20}".format("increment", "simulation time", "inc compute time"))
                for k in timingTypes:
                    f.write(" {:<20}".format(k))
                f.write("\n#\n")

    def finalizeIncrement(self, statusInfoDict=None, currentComputingTimes=None, **kwargs):
        self._record(statusInfoDict, currentComputingTimes)

    def finalizeFailedIncrement(self, statusInfoDict=None, currentComputingTimes=None, **kwargs):
        self._record(statusInfoDict, currentComputingTimes)

    def computeIncrementComputingTimes(self, currentComputingTimes):
        inc = createTimingDict()
        total = 0.0
        for k, v in currentComputingTimes.items():
            inc[k] = v - self.computingTimesOld[k]
            total += inc[k]
        return total, inc

    def _record(self, statusInfoDict, currentComputingTimes):
        d = statusInfoDict if statusInfoDict is not None else defaultdict(float)
        cur = currentComputingTimes if currentComputingTimes is not None else createTimingDict()
        total, inc = self.computeIncrementComputingTimes(cur)
        self.increments.append((int(d["inc"]), float(d["time end"]), total, inc))
        if self.journal is not None and total > 0.0:
            self.journal.printTable([("Time in {:}".format(k), " {:.5e}s ({:>4.1f}%)".format(v, v / total * 100))
                                     for k, v in inc.items()], self.identification)
        if self.exportFile is not None:
            with open(self.exportFile, "a") as f:
                f.write("  {:<20}".format(int(d["inc"])))
                f.write(" {:<20.5e}".format(d["time end"]))
                f.write(" {:<20.5e}".format(total))
                for v in inc.values():
                    f.write(" {:<20.5e}".format(v))
                f.write("\n")
        self.computingTimesOld = dict(cur)
