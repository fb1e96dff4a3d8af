"""Single-process communicator stand-in.

``op2.init`` takes ``MPI`` from the backend (reference: pyop2/op2.py:79-81).
This build runs on one GPU per process, so the communicator has rank 0 and
size 1 and reductions are the identity."""


class _SelfComm(object):
    rank = 0
    size = 1

    def Allreduce(self, sendbuf, recvbuf, op=None):
        recvbuf[...] = sendbuf


class MPIConfig(object):

    def __init__(self):
        self.COMM = _SelfComm()

    @property
    def parallel(self):
        """Are we running in parallel?"""
        return self.comm.size > 1

    @property
    def comm(self):
        """The communicator used by op2."""
        return self.COMM

    @comm.setter
    def comm(self, comm):
        if getattr(comm, 'size', 1) != 1:
            raise ValueError("only single-process communicators are supported")
        self.COMM = comm

    def rank_zero(self, f):
        """Decorator executing ``f`` only on rank zero."""
        def wrapper(*args, **kwargs):
            if self.comm.rank == 0:
                return f(*args, **kwargs)
        return wrapper


MPI = MPIConfig()
