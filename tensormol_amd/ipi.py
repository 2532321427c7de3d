"""i-PI socket clients: the drop-in one-structure client and a batched driver.

TMIPIManger(EnergyForceField, TCP_IP, TCP_PORT).md_run() has the call shape of
TensorMol/Interfaces/TMIPIinterface.py:7-61.  It speaks i-PI's driver protocol
as that client does:
  - every message starts with a 12-byte ASCII header, space padded;
  - STATUS -> the client answers READY, or HAVEDATA once it holds a result;
  - POSDATA is followed by the cell matrix (9 float64, bohr), its inverse
    (9 float64, 1/bohr), the atom count (int32) and the positions
    (3N float64, bohr);
  - GETFORCE -> FORCEREADY, then the energy (float64, Ha), the atom count
    (int32), the forces (3N float64, Ha/bohr), the virial (9 float64, zero
    here) and an extra string (int32 length, then its bytes).
EnergyForceField(positions in A) returns (E Ha, F J/mol/A), and the client
converts F with 1/(JOULEPERHARTREE * BOHRPERA), as the reference does.
Departures from the reference, which are bugs there: socket reads loop until a
message is complete (the reference takes whatever one recv returns), messages
are bytes (the reference sends str, which Python 3 rejects), an EXIT message
ends md_run instead of raising, and an INIT message (bead index, then a
length-prefixed string) is read and ignored instead of raising.

IPIBatchDriver serves several i-PI connections at once (the beads or replicas
of one i-PI simulation, each with its own client socket) and evaluates every
structure that has arrived in one packed engine launch, instead of one
evaluation per bead.
"""
import select
import socket

import numpy as np

from .params import BOHRPERA, JOULEPERHARTREE

HDRLEN = 12


def _header(msg):
    return msg.encode("ascii").ljust(HDRLEN)


class _Wire:
    """Exact-length reads and writes on a connected stream socket."""

    def __init__(self, sock):
        self.sock = sock

    def recv_exact(self, n):
        buf = bytearray()
        while len(buf) < n:
            chunk = self.sock.recv(n - len(buf))
            if not chunk:
                raise ConnectionError("i-PI server closed the connection")
            buf += chunk
        return bytes(buf)

    def header(self):
        return self.recv_exact(HDRLEN).decode("ascii", errors="replace").strip()

    def send(self, *parts):
        self.sock.sendall(b"".join(parts))

    def read_posdata(self):
        """(cell A, inverse cell 1/A, positions A (N, 3)) of a POSDATA message."""
        cell = np.frombuffer(self.recv_exact(72), dtype="<f8").reshape(3, 3) / BOHRPERA
        icell = np.frombuffer(self.recv_exact(72), dtype="<f8").reshape(3, 3) * BOHRPERA
        natom = int(np.frombuffer(self.recv_exact(4), dtype="<i4")[0])
        pos = np.frombuffer(self.recv_exact(24 * natom), dtype="<f8").reshape(natom, 3) / BOHRPERA
        return cell, icell, pos

    def read_init(self):
        self.recv_exact(4)                                   # bead index
        n = int(np.frombuffer(self.recv_exact(4), dtype="<i4")[0])
        self.recv_exact(n)

    def send_forces(self, energy, force_ha_bohr, extra=b"nothing"):
        f = np.ascontiguousarray(force_ha_bohr, dtype="<f8")
        self.send(_header("FORCEREADY"), np.float64(energy).astype("<f8").tobytes(),
                  np.int32(len(f)).astype("<i4").tobytes(), f.tobytes(),
                  np.zeros(9, dtype="<f8").tobytes(), np.int32(len(extra)).astype("<i4").tobytes(),
                  extra)


class TMIPIManger:
    """One i-PI client: each POSDATA is evaluated by EnergyForceField."""

    def __init__(self, EnergyForceField=None, TCP_IP="localhost", TCP_PORT=31415):
        self.EnergyForceField = EnergyForceField
        self.s = socket.socket(socket.AF_INET, socket.SOCK_STREAM)
        self.hasdata = False
        self.cell = None
        self.connected = False
        try:
            self.s.connect((TCP_IP, TCP_PORT))
            self.connected = True
        except OSError as e:
            print(f"Fail connect to server with address: {TCP_IP} {TCP_PORT} ({e})")

    def md_run(self):
        """Answer the server until it sends EXIT (or closes the socket)."""
        if not self.connected:
            raise ConnectionError("not connected to an i-PI server")
        w = _Wire(self.s)
        energy, force, natom = 0.0, None, 0
        try:
            while True:
                msg = w.header()
                if msg == "STATUS":
                    w.send(_header("HAVEDATA" if self.hasdata else "READY"))
                elif msg == "INIT":
                    w.read_init()
                elif msg == "POSDATA":
                    self.cell, _, position = w.read_posdata()
                    natom = len(position)
                    energy, force = self.EnergyForceField(position)
                    force = np.asarray(force, dtype=np.float64).reshape(natom, 3) \
                        / JOULEPERHARTREE / BOHRPERA
                    self.hasdata = True
                elif msg == "GETFORCE":
                    w.send_forces(float(energy), force)
                    self.hasdata = False
                elif msg == "EXIT":
                    return
                else:
                    raise Exception("wrong message from server")
        except ConnectionError:
            return
        finally:
            self.s.close()


class IPIBatchDriver:
    """Several i-PI client connections served by one engine.

    engine: a BPEEEngine; Z: atomic numbers of the simulated system (the same
    for every connection, as for the beads of one path-integral simulation);
    addresses: (host, port) per connection, or already connected sockets.
    periodic: evaluate under the minimum image of the (orthorhombic) cell the
    server sends.  Each round answers STATUS / GETFORCE and collects POSDATA on
    every connection that has a message; once every live connection holds a
    structure, or none has anything more to say within `wait` seconds, the
    structures are evaluated together in one packed launch.
    """

    def __init__(self, engine, Z, addresses, periodic=False, wait=0.05):
        self.eng = engine
        self.Z = np.asarray(Z, dtype=np.int64)
        self.periodic = periodic
        self.wait = wait
        self.wires = []
        for a in addresses:
            if isinstance(a, socket.socket):
                s = a
            else:
                s = socket.socket(socket.AF_INET, socket.SOCK_STREAM)
                s.connect(tuple(a))
            self.wires.append(_Wire(s))
        self._ws_key = None
        self._ws = None
        self.launches = 0
        self.evaluated = 0

    def _evaluate(self, items):
        """items: list of (wire, cell A, positions A) -> {wire: (E Ha, F Ha/bohr)}."""
        import torch
        n = len(self.Z)
        B = len(items)
        Zall = np.tile(self.Z, B)
        ms = np.arange(B + 1, dtype=np.int64) * n
        boxes = None
        if self.periodic:
            from .periodic import orthorhombic_edges
            boxes = np.stack([orthorhombic_edges(c) for _, c, _ in items])
        x = np.concatenate([p for _, _, p in items]).astype(np.float64)
        key = (B, None if boxes is None else boxes.tobytes())
        if key != self._ws_key:
            self._ws = self.eng.prepare(Zall, ms, boxes, pos0=x)
            self._ws_key = key
        pos = torch.tensor(x, dtype=torch.float64, device=self.eng.device)
        self.eng.run(self._ws, pos, check=True)
        r = self.eng.results(self._ws)
        self.launches += 1
        self.evaluated += B
        out = {}
        for k, (w, _, _) in enumerate(items):
            g = r["gradient"][ms[k]:ms[k + 1]]
            out[w] = (float(r["Etotal"][k]), -g / BOHRPERA)   # Ha/A -> force in Ha/bohr
        return out

    def run(self):
        """Serve until every connection has sent EXIT or closed."""
        live = list(self.wires)
        pending = {}       # wire -> (cell, positions) awaiting evaluation
        results = {}       # wire -> (E, F) awaiting GETFORCE
        try:
            while live:
                waiting = [w for w in live if w not in pending]
                ready = []
                if waiting:
                    ready, _, _ = select.select([w.sock for w in waiting], [], [],
                                                self.wait if pending else None)
                for sock in ready:
                    w = next(x for x in waiting if x.sock is sock)
                    try:
                        msg = w.header()
                    except ConnectionError:
                        live.remove(w)
                        continue
                    if msg == "STATUS":
                        w.send(_header("HAVEDATA" if w in results else "READY"))
                    elif msg == "INIT":
                        w.read_init()
                    elif msg == "POSDATA":
                        cell, _, p = w.read_posdata()
                        if len(p) != len(self.Z):
                            raise ValueError(f"i-PI sent {len(p)} atoms, the driver has {len(self.Z)}")
                        pending[w] = (cell, p)
                    elif msg == "GETFORCE":
                        e, f = results.pop(w)
                        w.send_forces(e, f)
                    elif msg == "EXIT":
                        live.remove(w)
                    else:
                        raise Exception("wrong message from server")
                if pending and (not ready or all(w in pending for w in live)):
                    items = [(w, c, p) for w, (c, p) in pending.items()]
                    results.update(self._evaluate(items))
                    pending.clear()
        finally:
            for w in self.wires:
                w.sock.close()
