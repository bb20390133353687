"""ensemble(simulation, parameters, keywords) (src/ensemble.jl:20-74) on one GPU.

The reference runs every member in an MPI communicator of its own.  Here the
members share the GPU: each member is a ``State`` with its own HIP stream,
driven by a host thread of its own (the library takes one host thread per
state, and ctypes releases the interpreter lock while a library call runs),
so the kernels of small members run side by side and fill the chip
together.  With ``batched`` (the default) the members' pressure solves run
together as well: a ``SolveBatch`` (pcf_ensemble) runs the BiCGSTAB
iterations of all solves in progress as one launch per kernel over the
members, each solve joining at an iteration boundary and leaving as soon as
it has converged.  Errors follow
reduce_exceptions: every member runs to its end or to its error, then the
first error is raised with the member named.
"""
from __future__ import annotations

import ctypes
import os
import threading
import time
from typing import Callable, Dict, List, Optional, Sequence, Tuple

from . import _abi
from .namelists import Namelists
from .state import PinCFlowError, State, integrate


class SolveBatch:
    """The pcf_ensemble of a set of states: their BiCGSTAB solves run in one
    batch, one launch per kernel of an iteration over all solves in progress
    (the line preconditioner in FP64 or FP32, x_size even, a multiple of 4 in
    the mixed solve; other solves run on their own); a solve joins and leaves
    the batch at iteration boundaries.  A state is attached from the thread
    that steps it and detached when that thread stops stepping it."""

    def __init__(self):
        self.lib = _abi.load_library()
        h = ctypes.c_void_p()
        rc = self.lib.pcf_ensemble_create(ctypes.byref(h))
        if rc:
            raise PinCFlowError(rc, self.lib.pcf_last_error().decode())
        self.h = h

    def attach(self, st: State):
        rc = self.lib.pcf_ensemble_attach(self.h, st.h)
        if rc:
            raise PinCFlowError(rc, self.lib.pcf_last_error().decode())

    def detach(self, st: State):
        self.lib.pcf_ensemble_detach(st.h)

    def stats(self) -> Dict[str, int]:
        out = (ctypes.c_int64 * 4)()
        self.lib.pcf_ensemble_stats(self.h, out)
        return {"batched_iterations": out[0], "member_solves": out[1], "solves_alone": out[2],
                "table_builds": out[3]}

    def close(self):
        if getattr(self, "h", None):
            rc = self.lib.pcf_ensemble_destroy(self.h)
            if rc:
                raise PinCFlowError(rc, self.lib.pcf_last_error().decode())
            self.h = None


class EnsembleError(RuntimeError):
    def __init__(self, member: int, err: BaseException):
        super().__init__(f"Ensemble member {member} has thrown the following exception: {err!r}")
        self.member = member
        self.error = err


def ensemble(simulation: Callable[..., Namelists], parameters: Dict[str, Sequence], keywords: Dict,
             delay: float = 0.0, max_concurrent: Optional[int] = None, write: bool = True,
             max_steps: int = 1 << 30, batched: bool = True, batch: Optional[SolveBatch] = None,
             on_step: Optional[Callable] = None) -> List[Tuple]:
    """Run ``integrate(simulation(**member_keywords))`` for every member.

    ``parameters`` maps keyword names to one value per member and must hold
    ``output_file`` (distinct per member, ensemble.jl:27-38); ``keywords``
    holds the keywords common to all members.  ``simulation`` returns the
    member's Namelists (the example constructors of ``examples`` do); the
    records of a member go to its output_file (``output.write_file``) when
    ``write``.  At most ``max_concurrent`` members (default: all) run at a
    time.  ``batched``: the members' solves run in batches (a SolveBatch of
    this call, or ``batch`` when given).  ``on_step(member, itime, info)``
    is called after every step of every member (from the member's thread).
    Returns, per member, integrate's (state, records, steps)."""
    if "output_file" not in parameters:
        raise ValueError("The parameter output_file must be assigned in ensembles!")
    files = list(parameters["output_file"])
    for entry in files:
        if files.count(entry) != 1:
            raise ValueError("There are duplicate output files!")
    n = len(files)
    for key, values in parameters.items():
        if len(values) != n:
            raise ValueError(f"parameter {key} has {len(values)} values for {n} members")
    results: List[Optional[Tuple]] = [None] * n
    errors: List[Optional[BaseException]] = [None] * n
    gate = threading.Semaphore(max_concurrent or n)
    own = batched and batch is None
    if own:
        batch = SolveBatch()
    elif not batched:
        batch = None

    def member(c: int):
        with gate:
            try:
                kw = {k: v[c] for k, v in parameters.items() if k != "output_file"}
                kw.update(keywords)
                nl = simulation(**kw)
                nl.output.output_file = files[c]
                if write and os.path.dirname(files[c]):
                    os.makedirs(os.path.dirname(files[c]), exist_ok=True)
                hook = (lambda itime, info: on_step(c + 1, itime, info)) if on_step else None
                results[c] = integrate(nl, write=write, max_steps=max_steps, batch=batch, on_step=hook)
            except BaseException as e:  # reported after every member has finished
                errors[c] = e

    threads = [threading.Thread(target=member, args=(c,), name=f"member-{c + 1}") for c in range(n)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if own:
        batch.close()
    for c, e in enumerate(errors):
        if e is not None:
            time.sleep(delay)
            raise EnsembleError(c + 1, e) from e
    return results
