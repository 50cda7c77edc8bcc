"""Host-side rules of the library that need no device, compiled from the
library's own headers with the host compiler."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "primal_dual_contacts_miror_amd", "csrc")

STALL_PROGRAM = r"""
#include <cstdio>
#include "pdc_stall.h"
using namespace pdc;
int main() {
	int bad = 0;
	auto expect = [&](StreamState s, double idle, StallAction want, const char* what) {
		if (stall_action(s, idle, kStallLimitSeconds) != want) { std::printf("FAIL %s\n", what); ++bad; }
	};
	// a failed stream ends the wait at once, whatever the time
	expect(StreamState::Failed, 0.0, StallAction::Fail, "failed stream, no idle time");
	expect(StreamState::Failed, 1e-3, StallAction::Fail, "failed stream, short idle time");
	// a drained stream without an outcome gets more launches
	expect(StreamState::Idle, 0.0, StallAction::QueueMore, "idle stream");
	expect(StreamState::Idle, 1e6, StallAction::QueueMore, "idle stream, long idle time");
	// a busy stream is waited for, up to the limit
	expect(StreamState::Busy, 0.0, StallAction::Wait, "busy stream");
	expect(StreamState::Busy, kStallLimitSeconds * 0.99, StallAction::Wait, "busy stream below the limit");
	expect(StreamState::Busy, kStallLimitSeconds * 1.01, StallAction::Fail, "busy stream past the limit");
	std::printf("%d failures\n", bad);
	return bad;
}
"""


def test_stall_rule_of_the_pcg_wait(tmp_path):
    """The asynchronous PCG loop polls progress records written by the
    kernels.  When the device position stops moving, a stream error must end
    the call (not spin forever), a drained stream gets more launches, and a
    busy stream is waited for only up to a time limit."""
    src = tmp_path / "stall.cpp"
    src.write_text(STALL_PROGRAM)
    exe = tmp_path / "stall"
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-Wall", "-I", CSRC, str(src), "-o", str(exe)])
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
