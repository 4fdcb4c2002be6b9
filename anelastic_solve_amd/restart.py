"""Restart dumps: the reference's serial flow file, read and written on the host.

A dump is a Fortran unformatted sequential file of eight records, in the
order of ``write_flow_small`` (serial branch, ``dns/Restart.F90:738-755``)
and ``Greadfile`` (``dns/Restart.F90:124-140``):

1. ``nstep`` (i4), ``time, delt, g, rho_0, scal1_0, Re, Pr, Sc`` (8 x r8)
2. ``nxp2, nyp2, nzp2`` (3 x i4): the ghosted extents n+2
3. ``u``, 4. ``v``, 5. ``w``, 6. ``p``, 7. ``rho``, 8. ``scal1``: each the
   whole ghosted field, r8, column-major (x fastest)

Every record is framed by its byte length as a 4-byte integer before and
after the payload (the sequential-file layout of the Fortran compilers the
reference builds with).  The byte order is detected from the first marker,
so dumps written big-endian (``-convert big_endian``) read too.  Records of
2 GiB and more, which those compilers split into sub-records, are refused.

Fields come back as numpy arrays of shape ``(nz+2, ny+2, nx+2)``: the
reference's memory order, the same layout ``mg_solver``, ``Psource``,
``Pgrad`` and ``divergence`` take.  File names follow ``concat``/``concati``
(``grid/utils/string.F90:1-77``): the trimmed base name followed by the
step number written ``(i7)`` and left-packed.
"""
from dataclasses import dataclass, field
import os

import numpy as np

FIELDS = ("u", "v", "w", "p", "rho", "scal1")
HEADER = ("nstep", "time", "delt", "g", "rho_0", "scal1_0", "Re", "Pr", "Sc")
_HEADER_BYTES = 4 + 8 * 8
_MAX_RECORD = 2 ** 31 - 1


@dataclass
class FlowDump:
    """One restart dump: the header scalars and the six ghosted fields."""
    nstep: int = 0
    time: float = 0.0
    delt: float = 0.0
    g: float = 0.0
    rho_0: float = 0.0
    scal1_0: float = 0.0
    Re: float = 0.0
    Pr: float = 0.0
    Sc: float = 0.0
    fields: dict = field(default_factory=dict)

    @property
    def shape(self):
        """Ghosted shape ``(nz+2, ny+2, nx+2)``."""
        return self.fields["u"].shape

    def __getattr__(self, name):
        if name in FIELDS and "fields" in self.__dict__ and name in self.__dict__["fields"]:
            return self.__dict__["fields"][name]
        raise AttributeError(name)

    def to_device(self, names=FIELDS, device="cuda"):
        """The named fields as FP64 torch tensors on ``device`` (one copy each)."""
        import torch
        return {n: torch.from_numpy(np.ascontiguousarray(self.fields[n])).to(device) for n in names}


def restart_name(basename, n):
    """``concat(ICfile, basename)`` + ``concati(ICfile, n)``: the name with its
    blanks trimmed, then ``n`` as ``(i7)`` left-packed (asterisks when it
    does not fit in 7 characters, as the Fortran edit descriptor writes)."""
    digits = "%d" % n
    if len(digits) > 7:
        digits = "*" * 7
    return basename.strip() + digits


def _order(marker_bytes):
    le = int.from_bytes(marker_bytes, "little", signed=True)
    be = int.from_bytes(marker_bytes, "big", signed=True)
    if le == _HEADER_BYTES:
        return "<"
    if be == _HEADER_BYTES:
        return ">"
    raise ValueError("not a restart dump: the first record is %d bytes, the header is %d"
                     % (le, _HEADER_BYTES))


def _read_record(f, bo, nbytes, what):
    m = f.read(4)
    if len(m) < 4:
        raise ValueError("restart dump ends before the %s record" % what)
    n = int.from_bytes(m, "little" if bo == "<" else "big", signed=True)
    if n < 0:
        raise ValueError("the %s record is split into sub-records (2 GiB or more): not supported" % what)
    if n != nbytes:
        raise ValueError("the %s record holds %d bytes, expected %d" % (what, n, nbytes))
    data = f.read(n)
    if len(data) < n:
        raise ValueError("restart dump ends inside the %s record" % what)
    t = f.read(4)
    if len(t) < 4 or int.from_bytes(t, "little" if bo == "<" else "big", signed=True) != n:
        raise ValueError("the %s record's end marker does not match its length" % what)
    return data


def read_dump(path, shape=None):
    """Read one dump file (``Greadfile``, ``dns/Restart.F90:86-159``).

    ``shape`` is the run's ghosted ``(nz+2, ny+2, nx+2)``; when it is given a
    dump of another size is an error, as in ``Greadfile:132-134``.
    """
    with open(path, "rb") as f:
        head = f.read(4)
        if len(head) < 4:
            raise ValueError("restart dump %s is empty" % path)
        bo = _order(head)
        f.seek(0)
        h = _read_record(f, bo, _HEADER_BYTES, "header")
        nstep = int(np.frombuffer(h[:4], dtype=bo + "i4")[0])
        vals = np.frombuffer(h[4:], dtype=bo + "f8")
        d = _read_record(f, bo, 12, "size")
        nxp2, nyp2, nzp2 = (int(x) for x in np.frombuffer(d, dtype=bo + "i4"))
        if min(nxp2, nyp2, nzp2) < 3:
            raise ValueError("restart dump sizes [%d %d %d] are not ghosted extents" % (nxp2, nyp2, nzp2))
        want = (nzp2, nyp2, nxp2)
        if shape is not None and tuple(shape) != want:
            raise ValueError("data file size is [nx,ny,nz] = [%d %d %d]; run size is [%d %d %d]"
                             % (nxp2, nyp2, nzp2, shape[2], shape[1], shape[0]))
        nb = 8 * nxp2 * nyp2 * nzp2
        if nb > _MAX_RECORD:
            raise ValueError("a field of %d bytes needs sub-records (2 GiB or more): not supported" % nb)
        fields = {}
        for name in FIELDS:
            raw = _read_record(f, bo, nb, name)
            a = np.frombuffer(raw, dtype=bo + "f8").reshape(want)
            fields[name] = a.astype("<f8") if bo == ">" else a.copy()
    dump = FlowDump(nstep, *(float(x) for x in vals), fields=fields)
    return dump


def write_dump(path, dump, byteorder="<"):
    """Write one dump file with the record sequence of ``write_flow_small``
    (``dns/Restart.F90:738-755``).  Existing files are replaced
    (``status='unknown'``)."""
    shape = dump.shape
    for n in FIELDS:
        if dump.fields[n].shape != shape:
            raise ValueError("field %s has shape %s, u has %s" % (n, dump.fields[n].shape, shape))
    nzp2, nyp2, nxp2 = shape
    nb = 8 * nxp2 * nyp2 * nzp2
    if nb > _MAX_RECORD:
        raise ValueError("a field of %d bytes needs sub-records (2 GiB or more): not supported" % nb)
    bo = byteorder
    i4, f8 = bo + "i4", bo + "f8"

    def rec(f, payload):
        m = np.array([len(payload)], dtype=i4).tobytes()
        f.write(m)
        f.write(payload)
        f.write(m)

    tmp = str(path) + ".tmp"
    with open(tmp, "wb") as f:
        hdr = np.array([dump.nstep], dtype=i4).tobytes() + np.array(
            [dump.time, dump.delt, dump.g, dump.rho_0, dump.scal1_0, dump.Re, dump.Pr, dump.Sc],
            dtype=f8).tobytes()
        rec(f, hdr)
        rec(f, np.array([nxp2, nyp2, nzp2], dtype=i4).tobytes())
        for n in FIELDS:
            rec(f, np.ascontiguousarray(dump.fields[n], dtype=f8).tobytes())
    os.replace(tmp, path)


def read_flow(in_file, in_iter, shape=None):
    """``read_flow(in_file, in_iter, stat)`` (``dns/Restart.F90:1-42``), serial
    branch: reads the dump ``in_file`` + ``in_iter`` (``start_small``,
    ``:259-464``)."""
    return read_dump(restart_name(in_file, in_iter), shape)


def write_flow(out_file, dump, flowDIR=""):
    """``write_flow(out_file, stat)`` (``dns/Restart.F90:44-84``), serial
    branch: writes ``flowDIR`` + ``out_file`` + ``nstep``; returns the path."""
    path = restart_name(flowDIR + out_file, dump.nstep)
    write_dump(path, dump)
    return path
