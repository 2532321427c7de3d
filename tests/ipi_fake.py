"""A minimal i-PI server for the client tests: it drives one or more client
connections through the driver protocol (STATUS / POSDATA / GETFORCE / EXIT)
for a list of structures and records what the clients return.  Messages are
sent in small chunks, so a client that takes one recv() as a whole message
fails."""
import socket
import threading
import time

import numpy as np

BOHRPERA = 1.889725989


def _hdr(m):
    return m.encode().ljust(12)


def _recv(c, n):
    buf = b""
    while len(buf) < n:
        chunk = c.recv(n - len(buf))
        if not chunk:
            raise ConnectionError("client closed")
        buf += chunk
    return buf


def _send_chunked(c, data, step=7):
    for k in range(0, len(data), step):
        c.sendall(data[k:k + step])
        time.sleep(0.0005)


class FakeIPIServer:
    """frames[k][c]: positions (A) sent to connection c at step k; cell (A)."""

    def __init__(self, frames, cell, nconn=1):
        self.frames = frames
        self.cell = np.asarray(cell, dtype=np.float64)
        self.nconn = nconn
        self.srv = socket.socket(socket.AF_INET, socket.SOCK_STREAM)
        self.srv.setsockopt(socket.SOL_SOCKET, socket.SO_REUSEADDR, 1)
        self.srv.bind(("127.0.0.1", 0))
        self.srv.listen(nconn)
        self.port = self.srv.getsockname()[1]
        self.results = [[None] * nconn for _ in frames]
        self.replies = []
        self.error = None
        self.thread = threading.Thread(target=self._serve, daemon=True)

    def start(self):
        self.thread.start()
        return self

    def _serve(self):
        try:
            conns = [self.srv.accept()[0] for _ in range(self.nconn)]
            cellb = (self.cell * BOHRPERA).astype("<f8").tobytes()
            icellb = (np.linalg.inv(self.cell) / BOHRPERA).astype("<f8").tobytes()
            for k, frame in enumerate(self.frames):
                for c, pos in zip(conns, frame):
                    c.sendall(_hdr("STATUS"))
                    self.replies.append(_recv(c, 12).decode().strip())
                    msg = (_hdr("POSDATA") + cellb + icellb +
                           np.int32(len(pos)).astype("<i4").tobytes() +
                           (np.asarray(pos) * BOHRPERA).astype("<f8").tobytes())
                    _send_chunked(c, msg)
                for ci, (c, pos) in enumerate(zip(conns, frame)):
                    c.sendall(_hdr("STATUS"))
                    self.replies.append(_recv(c, 12).decode().strip())
                    c.sendall(_hdr("GETFORCE"))
                    head = _recv(c, 12).decode().strip()
                    e = float(np.frombuffer(_recv(c, 8), "<f8")[0])
                    n = int(np.frombuffer(_recv(c, 4), "<i4")[0])
                    f = np.frombuffer(_recv(c, 24 * n), "<f8").reshape(n, 3).copy()
                    vir = np.frombuffer(_recv(c, 72), "<f8")
                    ln = int(np.frombuffer(_recv(c, 4), "<i4")[0])
                    extra = _recv(c, ln)
                    self.results[k][ci] = dict(head=head, energy=e, forces=f, virial=vir, extra=extra)
            for c in conns:
                c.sendall(_hdr("EXIT"))
                c.close()
        except Exception as ex:   # reported by the test
            self.error = ex
        finally:
            self.srv.close()

    def join(self, timeout=60):
        self.thread.join(timeout)
        assert not self.thread.is_alive(), "fake i-PI server did not finish"
        if self.error is not None:
            raise self.error
