// pdc_host.h — C++ host mirror of the reference's Contact/ and Solver/
// interfaces, implemented over the C ABI of libpdc (include/pdc.h).
//
//   ContactInformation       src/Contact/ContactInfo.h:12-145 (same fields and J operators)
//   RigidObject              the RigidObjectBase state the solver reads/writes
//                            (src/CollisionObject/RigidObjectBase.h:114-198)
//   ForceSolver              src/Solver/ForceSolverBase.h:43-97
//   GpuPrimalDualForceSolver PrimalDualForceSolver (src/Solver/PrimalDual/PrimalDualForceSolver.h:12-164)
//                            with the device doing fillForces
//   DynamicSystem            DynamicSystem::step / contactDetection (src/Dynamics/DynamicSystem.cpp:106-218)
//
// Configuration errors throw std::runtime_error, as the reference does; a
// failing device call throws std::runtime_error carrying pdc_last_error().
#pragma once

#include <array>
#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pdc.h"

namespace pdc_host {

struct ContactInformation {
	double pos[3] = {0, 0, 0};         // global collision position
	double localForce[3] = {0, 0, 0};  // contact force in the local frame
	double depth = 0;                  // penetration depth
	size_t objId[2] = {0, 0};          // contacting objects
	double transformation[2][18] = {}; // (negative) constraint Jacobians, 6x3 column-major

	// normal(): first column of the side-0 frame
	std::array<double, 3> normal() const {
		return {transformation[0][0], transformation[0][1], transformation[0][2]};
	}
	// u = -T0^T v_a - T1^T v_b, v a 6-per-body velocity vector
	std::array<double, 3> toLocal(const double* vels) const;
	// -T_j * local
	std::array<double, 6> toGlobal(int j, const double* local) const;
};

struct RigidObject {
	int shape = PDC_SHAPE_BOX;
	double dims[3] = {0.5, 0.5, 0.5};  // box half extents / sphere radius in dims[0]
	double mass = 1;
	double inertia[3] = {1, 1, 1};     // principal moments, body frame
	bool isStatic = false;
	bool emitSamples = true;
	double translation[3] = {0, 0, 0};
	double rotation[4] = {0, 0, 0, 1}; // x y z w
	double velocity[3] = {0, 0, 0};
	double angularVel[3] = {0, 0, 0};
	double acceleration[3] = {0, 0, 0};  // velocity increment written by fillForces
	double angularAcc[3] = {0, 0, 0};
	bool isActive() const { return !isStatic; }
};

enum ForceSolverType { GaussSeidel, PrimalDual };

class ForceSolver {
   public:
	virtual ~ForceSolver() {}
	virtual ForceSolverType getType() = 0;
	virtual void reset() = 0;
	virtual void fillForces(double timestep, const double gravity[3],
	                        std::vector<ContactInformation>& contactInfo,
	                        std::vector<std::unique_ptr<RigidObject>>& objects) = 0;
	virtual std::string printTimings(size_t totalSteps) = 0;
};

// Owns one pdc_context.  Bodies are uploaded on first use and whenever the
// object table changes size.
class GpuPrimalDualForceSolver : public ForceSolver {
   public:
	explicit GpuPrimalDualForceSolver(const pdc_params& params, int device = 0);
	~GpuPrimalDualForceSolver() override;
	GpuPrimalDualForceSolver(const GpuPrimalDualForceSolver&) = delete;
	GpuPrimalDualForceSolver& operator=(const GpuPrimalDualForceSolver&) = delete;

	ForceSolverType getType() override { return PrimalDual; }
	void reset() override;
	void fillForces(double timestep, const double gravity[3],
	                std::vector<ContactInformation>& contactInfo,
	                std::vector<std::unique_ptr<RigidObject>>& objects) override;
	std::string printTimings(size_t totalSteps) override;

	void configure(const pdc_params& params);
	const pdc_params& params() const { return m_params; }
	const pdc_solve_stats& lastStats() const { return m_last; }
	pdc_context* context() { return m_ctx; }
	void uploadBodies(const std::vector<std::unique_ptr<RigidObject>>& objects);

   private:
	void check(int rc, const char* what);
	pdc_context* m_ctx = nullptr;
	pdc_params m_params;
	size_t m_uploaded = (size_t)-1;
	pdc_solve_stats m_last{};
	size_t m_totalIterations = 0;
};

// World with device-resident state: contacts and solve run on the device,
// the host copy of the objects is refreshed by syncState().
class DynamicSystem {
   public:
	explicit DynamicSystem(const pdc_params& params, int device = 0);
	void addObject(std::unique_ptr<RigidObject> obj) { m_objects.push_back(std::move(obj)); }
	void initialize();                  // uploads bodies and state
	void step();                        // one timestep on the device
	void contactDetection();            // fills contacts() from the device
	void syncState();                   // device state -> objects
	double time() const { return m_time; }
	size_t nObjects() const { return m_objects.size(); }
	RigidObject& getObject(size_t i) { return *m_objects[i]; }
	std::vector<ContactInformation>& contacts() { return m_contactInfo; }
	GpuPrimalDualForceSolver& solver() { return m_solver; }
	std::string printTimings() const;

   private:
	GpuPrimalDualForceSolver m_solver;
	std::vector<std::unique_ptr<RigidObject>> m_objects;
	std::vector<ContactInformation> m_contactInfo;
	double m_time = 0;
	size_t m_steps = 0;
};

}  // namespace pdc_host
