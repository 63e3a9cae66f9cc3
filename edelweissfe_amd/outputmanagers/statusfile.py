"""Status file: one line per increment, converged or failed (statusfile.py:46-124).

Columns step, inc, iters, converged, time inc, time end and the notes (the message of the
exception that failed the increment).  A failed increment reports ``iters`` as inf, as the
reference solver sets it (nonlinearimplicitstatic.py:252, :269).  The header is written
once per definition, before the first line; the file name defaults to ``job.sta``.  The
column widths and number formats are those of the reference, so that status files of both
implementations compare line by line."""

from .base import OutputManagerBase

_HEADER = "#{: >5}{: >6}{: >6}{: >10}{: >12}{: >12}    {:<}\n"
_LINE = "{: >6}{: >6}{: >6}{: >10}{: >12.3e}{: >12.3e}    # {:<s}\n"


class OutputManager(OutputManagerBase):
    identification = "Statusfile"

    def __init__(self, name=None, model=None, fieldOutputController=None, journal=None, plotter=None):
        self.journal = journal
        self.filename = "job.sta"
        self.statusFileExists = False

    def updateDefinition(self, **kwargs):
        if "filename" in kwargs:
            self.filename = kwargs["filename"]
        self.statusFileExists = False

    def finalizeIncrement(self, statusInfoDict=None, **kwargs):
        self.writeStatusFile(statusInfoDict or {})

    def finalizeFailedIncrement(self, statusInfoDict=None, **kwargs):
        self.writeStatusFile(statusInfoDict or {})

    def writeStatusFile(self, statusInfoDict):
        d = statusInfoDict
        if not self.statusFileExists:
            with open(self.filename, "w") as f:
                f.write("#\n# This is a status file of EdelweissFE.\n#\n")
                f.write(_HEADER.format("step", "inc", "iters", "converged", "time inc", "time end", "notes"))
                f.write("#\n")
            self.statusFileExists = True
        with open(self.filename, "a") as f:
            f.write(_LINE.format(d["step"], d["inc"], d["iters"], d["converged"], d["time inc"], d["time end"],
                                 d["notes"]))


def read_status_file(filename):
    """Rows of a status file as dicts (step, inc, iters, converged, time inc, time end, notes);
    iters is a float (inf for a failed increment), converged a bool."""
    rows = []
    with open(filename) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            head, _, notes = line.partition("    # ")
            step, inc, iters, conv, tinc, tend = head.split()
            rows.append({"step": int(step), "inc": int(inc), "iters": float(iters),
                         "converged": conv in ("True", "1"), "time inc": float(tinc), "time end": float(tend),
                         "notes": notes.rstrip("\n")})
    return rows
