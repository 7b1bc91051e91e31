"""msgpack-RPC server (CHANS-HiPIMS_server.py:269-340 semantics) over a real
127.0.0.1 socket, driven the way the FLAMEGPU client drives it
(functions.c:85-99): request_data(t), then set_target(t + 120).  The model is
a stand-in with a fixed step so the run-ahead and blocking rules are exact."""
import threading
import time

import numpy as np
import pytest

from hipims_amd.server import TIME_STEP, FloodServer, RPCClient, serve


class FakeSimulation:
    def __init__(self, dt=7.0, delay=0.0):
        self.t = 0.0
        self.dt = dt
        self.delay = delay
        self.steps = 0
        self.results = None
        self.resets = 0

    def simulate_one_time_period(self):
        time.sleep(self.delay)
        self.t += self.dt
        self.steps += 1
        self.results = np.array([self.t / 1000.0, 2 * self.t / 1000.0])

    def time(self):
        return self.t

    def get_latest_data(self):
        return self.results

    def reset_simulation(self):
        self.resets += 1


@pytest.fixture()
def running():
    sim = FakeSimulation()
    flood = FloodServer(sim)
    srv = serve(flood, "127.0.0.1", 0)
    yield sim, flood, srv.server_address[1]
    srv.shutdown()
    srv.server_close()


def test_client_sequence(running):
    sim, flood, port = running
    c = RPCClient("127.0.0.1", port)
    assert c.call("echo", "ping") == "ping"
    assert c.call("echo", [1, 2.5]) == [1, 2.5]
    for k in range(1, 4):
        t = k * TIME_STEP
        data = c.call("request_data", t)
        # the run stops at the first step reaching t; current time is truncated to int
        assert sim.t >= t and sim.t - sim.dt < t
        assert flood.current_time == int(sim.t)
        assert data == pytest.approx([sim.t / 1000.0, 2 * sim.t / 1000.0], rel=0, abs=0)
        mm = [int(np.int16(v * 1000)) for v in data]  # the client's short(h * 1000)
        assert mm[0] == int(sim.t)
        c.call("set_target", t + TIME_STEP)  # runs ahead in the background
    c.close()


def test_wrong_time_and_reset(running):
    sim, flood, port = running
    c = RPCClient("127.0.0.1", port)
    c.call("request_data", 240)
    with pytest.raises(RuntimeError, match="Wrong Time"):
        c.call("set_target", 100)
    c.call("reset_simulation")
    assert flood.current_time == 0 and flood.target_time == 0 and sim.resets == 1
    # the model state is kept (as in the reference): the next request steps once
    t_before = sim.t
    c.call("request_data", 10)
    assert sim.t == t_before + sim.dt
    with pytest.raises(RuntimeError, match="method not found"):
        c.call("no_such_method")
    c.close()


def test_requests_wait_for_a_running_target():
    sim = FakeSimulation(dt=10.0, delay=0.002)
    flood = FloodServer(sim)
    srv = serve(flood, "127.0.0.1", 0)
    try:
        c = RPCClient("127.0.0.1", srv.server_address[1])
        c.call("set_target", 600)  # returns at once; 60 steps in the worker
        data = c.call("get_latest_data")  # waits for the run to finish
        assert sim.t == 600.0 and data[0] == pytest.approx(0.6)
        # two clients at once: both see the finished run
        out = []
        th = [threading.Thread(target=lambda: out.append(RPCClient("127.0.0.1", srv.server_address[1])
                                                         .call("request_data", 900))) for _ in range(2)]
        for t in th:
            t.start()
        for t in th:
            t.join(30)
        assert len(out) == 2 and out[0] == out[1] and sim.t == 900.0
    finally:
        srv.shutdown()
        srv.server_close()
