// pdc_stall.h — what the host does while it waits on the progress records of
// a running solve and the device's position has stopped moving.  Free of HIP
// types, so that the rule can be compiled and tested on a host alone.
#pragma once

namespace pdc {

// State of the context's stream, as its query reports it.
enum class StreamState { Idle, Busy, Failed };

enum class StallAction {
	Wait,       // the stream is still working: keep polling
	QueueMore,  // the stream drained before the solve reported an outcome: queue more launches
	Fail        // the stream reported an error, or nothing moved for too long: end the call with an error
};

// `idle_seconds`: time since the device's position last moved.
inline StallAction stall_action(StreamState s, double idle_seconds, double limit_seconds) {
	if (s == StreamState::Failed) return StallAction::Fail;
	if (s == StreamState::Idle) return StallAction::QueueMore;
	return idle_seconds > limit_seconds ? StallAction::Fail : StallAction::Wait;
}

// A solve whose device position has not moved for this long is taken as
// lost (one PCG iteration takes microseconds, a phase of a large batch
// milliseconds).
constexpr double kStallLimitSeconds = 60.0;

}  // namespace pdc
