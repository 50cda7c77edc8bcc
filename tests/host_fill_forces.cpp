// host_fill_forces.cpp — test driver of the C++ host mirror: reads a body
// table, their state and a contact list from a binary file, runs
// pdc_host::GpuPrimalDualForceSolver::fillForces (the ForceSolver plugin
// over the C ABI) and writes the velocity increments, the contacts' local
// forces, the solve statistics and whether every contact's objId came back
// unchanged (the reference's fillForces renumbers and restores them).
//
// Input (native byte order): pdc_params; int32 N, C; int32 shape[N],
// is_static[N], emit[N]; double dims[3N], mass[N], inertia[3N], quat[4N]
// (x y z w), vel[3N], angvel[3N]; int32 ids[2C]; double T[36C], depth[C].
// Output: double acc[6N], local_force[3C]; pdc_solve_stats; int32 ids_kept.
// Exit status: 0 ok, 1 bad input, 2 the solver threw.
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "../primal_dual_contacts_miror_amd/host/pdc_host.h"

template <class T>
static bool get(std::FILE* f, T* p, size_t n) {
	return std::fread(p, sizeof(T), n, f) == n;
}

int main(int argc, char** argv) {
	if (argc != 3) {
		std::fprintf(stderr, "usage: %s input output\n", argv[0]);
		return 1;
	}
	std::FILE* in = std::fopen(argv[1], "rb");
	if (!in) return 1;
	pdc_params p;
	int32_t N = 0, C = 0;
	if (!get(in, &p, 1) || !get(in, &N, 1) || !get(in, &C, 1) || N < 0 || C < 0) return 1;
	std::vector<int32_t> shape(N), stat(N), emit(N), ids(2 * (size_t)C);
	std::vector<double> dims(3 * (size_t)N), mass(N), inertia(3 * (size_t)N), quat(4 * (size_t)N),
	    vel(3 * (size_t)N), ang(3 * (size_t)N), T(36 * (size_t)C), depth(C);
	bool ok = get(in, shape.data(), N) && get(in, stat.data(), N) && get(in, emit.data(), N) &&
	          get(in, dims.data(), 3 * (size_t)N) && get(in, mass.data(), N) &&
	          get(in, inertia.data(), 3 * (size_t)N) && get(in, quat.data(), 4 * (size_t)N) &&
	          get(in, vel.data(), 3 * (size_t)N) && get(in, ang.data(), 3 * (size_t)N) &&
	          get(in, ids.data(), 2 * (size_t)C) && get(in, T.data(), 36 * (size_t)C) && get(in, depth.data(), C);
	std::fclose(in);
	if (!ok) return 1;
	std::vector<std::unique_ptr<pdc_host::RigidObject>> objects;
	for (int32_t i = 0; i < N; ++i) {
		auto o = std::make_unique<pdc_host::RigidObject>();
		o->shape = shape[i];
		std::memcpy(o->dims, &dims[3 * i], sizeof(o->dims));
		o->mass = mass[i];
		std::memcpy(o->inertia, &inertia[3 * i], sizeof(o->inertia));
		o->isStatic = stat[i] != 0;
		o->emitSamples = emit[i] != 0;
		std::memcpy(o->rotation, &quat[4 * i], sizeof(o->rotation));
		std::memcpy(o->velocity, &vel[3 * i], sizeof(o->velocity));
		std::memcpy(o->angularVel, &ang[3 * i], sizeof(o->angularVel));
		objects.push_back(std::move(o));
	}
	std::vector<pdc_host::ContactInformation> contacts(C);
	for (int32_t k = 0; k < C; ++k) {
		contacts[k].objId[0] = (size_t)ids[2 * k];
		contacts[k].objId[1] = (size_t)ids[2 * k + 1];
		std::memcpy(contacts[k].transformation, &T[36 * (size_t)k], 36 * sizeof(double));
		contacts[k].depth = depth[k];
	}
	std::vector<double> acc(6 * (size_t)N), lf(3 * (size_t)C);
	pdc_solve_stats st{};
	int32_t kept = 1;
	try {
		pdc_host::GpuPrimalDualForceSolver solver(p);
		solver.fillForces(p.timestep, p.gravity, contacts, objects);
		st = solver.lastStats();
	} catch (const std::exception& e) {
		std::fprintf(stderr, "fillForces: %s\n", e.what());
		return 2;
	}
	for (int32_t i = 0; i < N; ++i) {
		std::memcpy(&acc[6 * i], objects[i]->acceleration, 3 * sizeof(double));
		std::memcpy(&acc[6 * i + 3], objects[i]->angularAcc, 3 * sizeof(double));
	}
	for (int32_t k = 0; k < C; ++k) {
		std::memcpy(&lf[3 * k], contacts[k].localForce, 3 * sizeof(double));
		if (contacts[k].objId[0] != (size_t)ids[2 * k] || contacts[k].objId[1] != (size_t)ids[2 * k + 1]) kept = 0;
	}
	std::FILE* out = std::fopen(argv[2], "wb");
	if (!out) return 1;
	std::fwrite(acc.data(), sizeof(double), acc.size(), out);
	std::fwrite(lf.data(), sizeof(double), lf.size(), out);
	std::fwrite(&st, sizeof(st), 1, out);
	std::fwrite(&kept, sizeof(kept), 1, out);
	std::fclose(out);
	return 0;
}
