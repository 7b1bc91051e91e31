"""msgpack-RPC flood server, wire-compatible with CHANS-HiPIMS_server.py.

The reference couples HiPIMS to the FLAMEGPU agent model through an RPC
server (CHANS-HiPIMS/CHANS-HiPIMS_server.py:269-340, gevent + mprpc) that an
rpclib client calls (CHANS-FLAMEGPU/.../src/model/functions.c:85-99): every
TIME_STEP = 120 model seconds the client calls ``request_data(t)`` (depths
at the agent-model cells, compact order), converts them to int16 millimetres
and calls ``set_target(t + 120)`` so the flood model runs ahead while the
agents move.  This module serves the same five methods over the same
msgpack-RPC framing ([0, id, method, params] -> [1, id, error, result];
[2, method, params] notifications) with the standard library's socketserver
and the ``msgpack`` package:

  request_data(time)       if time > current time: set_target(time); then
                           wait for the run to finish and return the depths
                           as a list of floats
  set_target(t, sync=True) "Wrong Time" error if t < current time; else run
                           the model in a worker thread until its t >= target
  get_latest_data()        wait, then the latest depths (a list here; the
                           reference returns a numpy array, which mprpc
                           cannot pack)
  reset_simulation()       wait, then set target and current time to 0 (the
                           model state is kept, as in the reference)
  echo(x)                  x

Semantics kept from the reference: the current time is the model's t after
each step and is truncated to int when a run ends (:280-286); requests wait
while a run is in progress (the reference spins on a flag, here a
condition variable).
"""
from __future__ import annotations

import socket
import socketserver
import threading

import msgpack
import numpy as np
import torch

TIME_STEP = 120  # functions.c:47


class Simulation:
    """The reference's ``Simulation`` (CHANS-HiPIMS_server.py:66-262) over a
    ``Godunov`` drop-in: one server step is addFlux, station rain, then the
    friction/update/time-control call; the result is h at the agent-model
    cells."""

    def __init__(self, numerical, rainfall_matrix, idx_target, device):
        self.numerical = numerical
        self.rainfallMatrix = np.asarray(rainfall_matrix, dtype=np.float64)
        self.idx_target = torch.as_tensor(np.asarray(idx_target, dtype=np.int64), device=device)
        self.device = device
        self.dt_list: list[float] = []
        self.t_list: list[float] = []
        self.results = None

    @staticmethod
    def abm_compact_index(dem_mask: np.ndarray, abm_mask: np.ndarray) -> np.ndarray:
        """Compact ids (row-major over the DEM's valid cells) of the cells the
        agent model covers (:139-153: the reference stamps those cells 9999
        in a copy of the DEM and keeps the stamped compact positions)."""
        valid = np.asarray(dem_mask) > 0
        return np.flatnonzero(np.asarray(abm_mask, bool)[valid]).astype(np.int64)

    def simulate_one_time_period(self):
        num = self.numerical
        num.addFlux()
        num.addStation_PrecipitationSource(self.rainfallMatrix, self.device)
        num.time_friction_euler_update_cuda(self.device)
        self.dt_list.append(num.dt.item())
        self.t_list.append(num.t.item())
        self.results = num.get_h()[self.idx_target]

    def time(self) -> float:
        return self.numerical.t.item()

    def get_latest_data(self):
        if self.results is None:
            raise RuntimeError("no step has run yet")
        return self.results.cpu()

    def reset_simulation(self):
        pass


class FloodServer:
    """The reference's ``Server`` methods (CHANS-HiPIMS_server.py:269-330)."""

    def __init__(self, simulation):
        self.simulation = simulation
        self.target_time = 0
        self.current_time = 0
        self._running = False
        self._cv = threading.Condition()
        self._worker: threading.Thread | None = None

    def _wait_idle(self):
        with self._cv:
            while self._running:
                self._cv.wait()

    def _run_loop(self):
        try:
            while self.current_time < self.target_time:
                self.simulation.simulate_one_time_period()
                self.current_time = self.simulation.time()
            self.current_time = int(self.current_time)
        finally:
            with self._cv:
                self._running = False
                self._cv.notify_all()

    def _sync_to_target_time(self):
        with self._cv:
            while self._running:
                self._cv.wait()
            self._running = True
        self._worker = threading.Thread(target=self._run_loop, daemon=True)
        self._worker.start()

    # -- RPC methods ------------------------------------------------------------
    def set_target(self, target_time, sync=True):
        target_time = int(target_time)
        if target_time < self.current_time:
            raise ValueError("Wrong Time")
        self.target_time = target_time
        if sync:
            self._sync_to_target_time()

    def request_data(self, time):
        time = int(time)
        if time > self.current_time:
            self.set_target(time)
        self._wait_idle()
        return self.get_latest_data()

    def get_latest_data(self):
        self._wait_idle()
        data = self.simulation.get_latest_data()
        if isinstance(data, torch.Tensor):
            data = data.numpy()
        return np.asarray(data).tolist()

    def reset_simulation(self):
        self._wait_idle()
        self.target_time = 0
        self.current_time = 0
        self.simulation.reset_simulation()

    def echo(self, x):
        return x

    METHODS = ("request_data", "set_target", "get_latest_data", "reset_simulation", "echo")

    def dispatch(self, method, params):
        if isinstance(method, bytes):
            method = method.decode()
        if method not in self.METHODS:
            raise AttributeError(f"method not found: {method}")
        return getattr(self, method)(*params)


class _Handler(socketserver.BaseRequestHandler):
    def handle(self):
        srv: FloodServer = self.server.flood
        unpacker = msgpack.Unpacker(raw=False)
        while True:
            try:
                chunk = self.request.recv(65536)
            except OSError:
                return
            if not chunk:
                return
            unpacker.feed(chunk)
            for msg in unpacker:
                if not isinstance(msg, (list, tuple)) or not msg:
                    continue
                if msg[0] == 0 and len(msg) == 4:
                    _, msgid, method, params = msg
                    try:
                        result, error = srv.dispatch(method, params), None
                    except Exception as e:  # reported to the caller, as mprpc does
                        result, error = None, str(e)
                    self.request.sendall(msgpack.packb([1, msgid, error, result], use_bin_type=True))
                elif msg[0] == 2 and len(msg) == 3:
                    try:
                        srv.dispatch(msg[1], msg[2])
                    except Exception:
                        pass


class _TCPServer(socketserver.ThreadingTCPServer):
    daemon_threads = True
    allow_reuse_address = True


def serve(flood: FloodServer, host: str = "127.0.0.1", port: int = 6060):
    """Start serving in a background thread; returns the socketserver (its
    ``server_address`` holds the bound port; ``shutdown()`` stops it)."""
    s = _TCPServer((host, port), _Handler)
    s.flood = flood
    threading.Thread(target=s.serve_forever, daemon=True).start()
    return s


class RPCClient:
    """Minimal msgpack-RPC client (the calls the rpclib client makes)."""

    def __init__(self, host: str = "127.0.0.1", port: int = 6060, timeout: float = 600.0):
        self.sock = socket.create_connection((host, port), timeout=timeout)
        self.unpacker = msgpack.Unpacker(raw=False)
        self.msgid = 0

    def call(self, method: str, *params):
        self.msgid += 1
        self.sock.sendall(msgpack.packb([0, self.msgid, method, list(params)], use_bin_type=True))
        while True:
            for msg in self.unpacker:
                if msg[0] == 1 and msg[1] == self.msgid:
                    if msg[2] is not None:
                        raise RuntimeError(msg[2])
                    return msg[3]
            chunk = self.sock.recv(65536)
            if not chunk:
                raise ConnectionError("server closed the connection")
            self.unpacker.feed(chunk)

    def close(self):
        self.sock.close()
