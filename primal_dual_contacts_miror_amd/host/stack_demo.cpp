// stack_demo.cpp — drives the C++ host mirror: a 10-cube stack on a ground
// box, stepped on the device; prints the reference's timing report.
// Exit status: 0 ok, 2 no usable device.
#include <cstdio>
#include <memory>

#include "pdc_host.h"

int main() {
	pdc_params p;
	pdc_default_params(&p);
	p.grid_n = 64;
	p.sample_spacing = 0.1;
	try {
		pdc_host::DynamicSystem ds(p);
		auto ground = std::make_unique<pdc_host::RigidObject>();
		ground->dims[0] = 10; ground->dims[1] = 0.5; ground->dims[2] = 10;
		ground->translation[1] = -0.5;
		ground->isStatic = true;
		ground->emitSamples = false;
		ds.addObject(std::move(ground));
		for (int k = 0; k < 10; ++k) {
			auto cube = std::make_unique<pdc_host::RigidObject>();
			cube->mass = 1;
			cube->inertia[0] = cube->inertia[1] = cube->inertia[2] = 1.0 / 6.0;
			cube->translation[1] = 0.5 + 0.001 + k * 1.001;
			ds.addObject(std::move(cube));
		}
		ds.initialize();
		for (int s = 0; s < 100; ++s) ds.step();
		ds.syncState();
		ds.contactDetection();
		std::printf("t=%.3f contacts=%zu top cube y=%.6f vy=%.3e\n", ds.time(), ds.contacts().size(),
		            ds.getObject(10).translation[1], ds.getObject(10).velocity[1]);
		std::printf("%s\n", ds.printTimings().c_str());
	} catch (const std::exception& e) {
		std::fprintf(stderr, "stack_demo: %s\n", e.what());
		return 2;
	}
	return 0;
}
