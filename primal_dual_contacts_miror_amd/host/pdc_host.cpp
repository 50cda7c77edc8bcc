// pdc_host.cpp — C++ host mirror of the reference's solver / dynamics
// interfaces over libpdc.  See pdc_host.h.
#include "pdc_host.h"

#include <cstring>
#include <sstream>

namespace pdc_host {

std::array<double, 3> ContactInformation::toLocal(const double* vels) const {
	// ContactInfo.h:46-51
	std::array<double, 3> u{};
	const double* va = vels + 6 * objId[0];
	const double* vb = vels + 6 * objId[1];
	for (int col = 0; col < 3; ++col) {
		double s0 = 0, s1 = 0;
		for (int r = 0; r < 6; ++r) {
			s0 += transformation[0][col * 6 + r] * va[r];
			s1 += transformation[1][col * 6 + r] * vb[r];
		}
		u[col] = -s0 - s1;
	}
	return u;
}

std::array<double, 6> ContactInformation::toGlobal(int j, const double* local) const {
	// ContactInfo.h:80-83
	std::array<double, 6> g{};
	for (int r = 0; r < 6; ++r) {
		double s = 0;
		for (int col = 0; col < 3; ++col) s += transformation[j][col * 6 + r] * local[col];
		g[r] = -s;
	}
	return g;
}

GpuPrimalDualForceSolver::GpuPrimalDualForceSolver(const pdc_params& params, int device)
	: m_params(params) {
	const int rc = pdc_create(device, &params, &m_ctx);
	if (rc == PDC_ERR_CONFIG) throw std::runtime_error("primal_dual solver: configuration rejected");
	if (rc != PDC_OK)
		throw std::runtime_error("primal_dual solver: no usable gfx950 device (status " +
		                         std::to_string(rc) + ")");
}

GpuPrimalDualForceSolver::~GpuPrimalDualForceSolver() { pdc_destroy(m_ctx); }

void GpuPrimalDualForceSolver::check(int rc, const char* what) {
	if (rc == PDC_OK) return;
	throw std::runtime_error(std::string(what) + ": " + pdc_last_error(m_ctx));
}

void GpuPrimalDualForceSolver::configure(const pdc_params& params) {
	check(pdc_configure(m_ctx, &params), "configure");
	m_params = params;
}

void GpuPrimalDualForceSolver::reset() {
	pdc_reset_timings(m_ctx);
	m_totalIterations = 0;
}

void GpuPrimalDualForceSolver::uploadBodies(const std::vector<std::unique_ptr<RigidObject>>& objects) {
	const size_t n = objects.size();
	std::vector<int32_t> shape(n), stat(n), emit(n);
	std::vector<double> dims(3 * n), mass(n), inertia(3 * n);
	for (size_t i = 0; i < n; ++i) {
		const RigidObject& o = *objects[i];
		shape[i] = o.shape;
		stat[i] = o.isStatic ? 1 : 0;
		emit[i] = o.emitSamples ? 1 : 0;
		mass[i] = o.mass;
		for (int a = 0; a < 3; ++a) {
			dims[3 * i + a] = o.dims[a];
			inertia[3 * i + a] = o.inertia[a];
		}
	}
	check(pdc_upload_bodies(m_ctx, (int32_t)n, shape.data(), dims.data(), mass.data(), inertia.data(),
	                        stat.data(), emit.data()),
	      "upload bodies");
	m_uploaded = n;
}

void GpuPrimalDualForceSolver::fillForces(double timestep, const double gravity[3],
                                          std::vector<ContactInformation>& contactInfo,
                                          std::vector<std::unique_ptr<RigidObject>>& objects) {
	if (timestep != m_params.timestep || gravity[0] != m_params.gravity[0] ||
	    gravity[1] != m_params.gravity[1] || gravity[2] != m_params.gravity[2]) {
		pdc_params p = m_params;
		p.timestep = timestep;
		std::memcpy(p.gravity, gravity, sizeof(p.gravity));
		configure(p);
	}
	if (m_uploaded != objects.size()) uploadBodies(objects);
	const size_t n = objects.size(), C = contactInfo.size();
	std::vector<double> quat(4 * n), vel(3 * n), ang(3 * n), acc(6 * n);
	for (size_t i = 0; i < n; ++i) {
		std::memcpy(&quat[4 * i], objects[i]->rotation, 4 * sizeof(double));
		std::memcpy(&vel[3 * i], objects[i]->velocity, 3 * sizeof(double));
		std::memcpy(&ang[3 * i], objects[i]->angularVel, 3 * sizeof(double));
	}
	std::vector<int32_t> ids(2 * C);
	std::vector<double> T(36 * C), depth(C), lf(3 * C);
	for (size_t i = 0; i < C; ++i) {
		ids[2 * i] = (int32_t)contactInfo[i].objId[0];
		ids[2 * i + 1] = (int32_t)contactInfo[i].objId[1];
		std::memcpy(&T[36 * i], contactInfo[i].transformation, 36 * sizeof(double));
		depth[i] = contactInfo[i].depth;
	}
	check(pdc_fill_forces(m_ctx, (int32_t)C, ids.data(), T.data(), depth.data(), quat.data(), vel.data(),
	                      ang.data(), acc.data(), lf.data(), &m_last),
	      "fillForces");
	for (size_t i = 0; i < n; ++i) {
		std::memcpy(objects[i]->acceleration, &acc[6 * i], 3 * sizeof(double));
		std::memcpy(objects[i]->angularAcc, &acc[6 * i + 3], 3 * sizeof(double));
	}
	for (size_t i = 0; i < C; ++i) std::memcpy(contactInfo[i].localForce, &lf[3 * i], 3 * sizeof(double));
	m_totalIterations += m_last.newton_iterations;
}

std::string GpuPrimalDualForceSolver::printTimings(size_t totalSteps) {
	// the fields of PrimalDualForceSolver::printTimings (PrimalDualForceSolver.cpp:636-648)
	pdc_timings t;
	pdc_get_timings(m_ctx, &t);
	const double s = totalSteps ? (double)totalSteps : 1.0;
	std::stringstream ss;
	ss << "\n\tAvg. force calc time per substep: " << t.force_process / s
	   << "\n\t\tBuild system: " << t.build_system / s << "\n\t\tSolve system: " << t.solve_system / s
	   << "\n\t\tLine search: " << t.line_search / s
	   << "\n\tAverage #iterations: " << (double)t.newton_iterations / s << "\n";
	return ss.str();
}

DynamicSystem::DynamicSystem(const pdc_params& params, int device) : m_solver(params, device) {}

void DynamicSystem::initialize() {
	m_solver.uploadBodies(m_objects);
	const size_t n = m_objects.size();
	std::vector<double> pos(3 * n), quat(4 * n), vel(3 * n), ang(3 * n);
	for (size_t i = 0; i < n; ++i) {
		std::memcpy(&pos[3 * i], m_objects[i]->translation, 3 * sizeof(double));
		std::memcpy(&quat[4 * i], m_objects[i]->rotation, 4 * sizeof(double));
		std::memcpy(&vel[3 * i], m_objects[i]->velocity, 3 * sizeof(double));
		std::memcpy(&ang[3 * i], m_objects[i]->angularVel, 3 * sizeof(double));
	}
	pdc_context* c = m_solver.context();
	if (pdc_set_state(c, pos.data(), quat.data(), vel.data(), ang.data()) != PDC_OK)
		throw std::runtime_error(std::string("set state: ") + pdc_last_error(c));
}

void DynamicSystem::step() {
	pdc_context* c = m_solver.context();
	pdc_solve_stats st;
	if (pdc_step(c, 1, &st) != PDC_OK) throw std::runtime_error(std::string("step: ") + pdc_last_error(c));
	m_time += m_solver.params().timestep;
	++m_steps;
}

void DynamicSystem::contactDetection() {
	pdc_context* c = m_solver.context();
	int32_t C = 0, ties = 0;
	if (pdc_detect_contacts(c, &C, &ties) != PDC_OK)
		throw std::runtime_error(std::string("contact detection: ") + pdc_last_error(c));
	std::vector<int32_t> ids(2 * (size_t)C);
	std::vector<double> pos(3 * (size_t)C), depth(C), T(36 * (size_t)C), lf(3 * (size_t)C);
	if (C > 0 && pdc_get_contacts(c, C, ids.data(), pos.data(), depth.data(), T.data(), lf.data()) != PDC_OK)
		throw std::runtime_error(std::string("get contacts: ") + pdc_last_error(c));
	m_contactInfo.assign(C, ContactInformation());
	for (int32_t i = 0; i < C; ++i) {
		ContactInformation& ci = m_contactInfo[i];
		ci.objId[0] = (size_t)ids[2 * i];
		ci.objId[1] = (size_t)ids[2 * i + 1];
		std::memcpy(ci.pos, &pos[3 * i], sizeof(ci.pos));
		ci.depth = depth[i];
		std::memcpy(ci.transformation, &T[36 * (size_t)i], sizeof(ci.transformation));
		std::memcpy(ci.localForce, &lf[3 * i], sizeof(ci.localForce));
	}
}

void DynamicSystem::syncState() {
	const size_t n = m_objects.size();
	std::vector<double> pos(3 * n), quat(4 * n), vel(3 * n), ang(3 * n);
	pdc_context* c = m_solver.context();
	if (pdc_get_state(c, pos.data(), quat.data(), vel.data(), ang.data()) != PDC_OK)
		throw std::runtime_error(std::string("get state: ") + pdc_last_error(c));
	for (size_t i = 0; i < n; ++i) {
		std::memcpy(m_objects[i]->translation, &pos[3 * i], 3 * sizeof(double));
		std::memcpy(m_objects[i]->rotation, &quat[4 * i], 4 * sizeof(double));
		std::memcpy(m_objects[i]->velocity, &vel[3 * i], 3 * sizeof(double));
		std::memcpy(m_objects[i]->angularVel, &ang[3 * i], 3 * sizeof(double));
	}
}

std::string DynamicSystem::printTimings() const {
	// the layout of DynamicSystem::printTimings (DynamicSystem.cpp:299-315)
	pdc_timings t;
	pdc_get_timings(const_cast<GpuPrimalDualForceSolver&>(m_solver).context(), &t);
	const double s = t.steps ? (double)t.steps : 1.0;
	std::stringstream ss;
	ss << "Times per substep (seconds):\n\tContact Detection: " << t.contact_detection / s
	   << "\n\tTotal time: " << (t.contact_detection + t.force_process) / s
	   << "\nAverage #contacts (after merge): " << (double)t.contacts / s
	   << "\nTotal #solids: " << m_objects.size() << "\n"
	   << const_cast<GpuPrimalDualForceSolver&>(m_solver).printTimings(t.steps);
	return ss.str();
}

}  // namespace pdc_host
