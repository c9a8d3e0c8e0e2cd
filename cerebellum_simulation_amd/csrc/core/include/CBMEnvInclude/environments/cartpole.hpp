// cartpole.hpp -- the cart-pole environment driven by two microzones (the
// interface of Matthew_Sim_CBM/includes/environments/cartpole.hpp:14-175,
// behaviour of src/environments/cartpole.cpp:24-285).
//
// The pole (half length 7.5, 0.1 on a cart of 1.0, g = 9.8) is integrated by
// Euler steps of 1 ms under the net force ForceRight - ForceLeft of microzones
// 1 and 0.  Five MF groups encode the state: 3% of the MFs fire at the excited
// rate always, and 6% each tile the pole velocity, pole angle, cart velocity
// and cart position with Gaussian tuning curves (updated only outside the
// 1500-step time-out that follows every fall).  Error goes to the microzone
// that should push against the pole's fall with probability min(|theta|, 0.01)
// per step; on a finite track also for cart position and velocity.  A trial
// ends when the pole passes +-10 degrees, the cart leaves the track, or the
// trial reaches maxTrialLen steps.
//
// One deliberate choice: the reference's Cartpole::step never calls the
// microzones' update(), so their forces keep their (uninitialised) starting
// value and the network's output never moves the cart.  closedLoop = false
// (the default, and what the command-line constructor gives) keeps that: the
// forces stay 0.  closedLoop = true updates both microzones from their zone's
// NC spikes at the start of every step, as Environment::step does for other
// environments.
#ifndef CBM_ENV_CARTPOLE_HPP
#define CBM_ENV_CARTPOLE_HPP

#include <cmath>
#include <fstream>
#include <limits>
#include <string>

#include "environment.hpp"

struct CartpoleOptions {
    std::string logfile;      // "" = no log
    int maxNumTrials = 20;
    long maxTrialLen = 1000000;
    float trackLen = 0;       // <= 0: infinite track
    std::string simStateDir;  // "" = no state files
    bool closedLoop = false;
    bool verbose = false;     // print one line per trial, as the reference does
};

class Cartpole : public Environment
{
public:
    Cartpole(CRandomSFMT0 *randGen, const CartpoleOptions &opts);
    // Command-line form: --logfile, --maxNumTrials, --maxTrialLen, --trackLen,
    // --simStateDir (defaults cartpole.log, 20, 1000000, 0, ./), --closedLoop;
    // unknown arguments are ignored.
    Cartpole(CRandomSFMT0 *randGen, int argc, char **argv);
    ~Cartpole();

    int numRequiredMZ() { return 2; }
    void setupMossyFibers(CBMState *simState);
    void setupMossyFibers(int numMF, int numNC);
    float *getState();
    void step(CBMSimCore *simCore);
    // The same step with the NC spikes of zones 0 and 1 given (closed loop)
    // rather than read from the core; errors still go to the core.
    void stepWithSpikes(CBMSimCore *simCore, const ct_uint8_t *apNCLeft, const ct_uint8_t *apNCRight);
    // The same with no core: errors are only counted by the microzones.
    void stepDetached(const ct_uint8_t *apNCLeft, const ct_uint8_t *apNCRight);
    bool terminated();
    // whether steps write the simulation state to files (simStateDir given)
    bool savesState() const { return !saveStateDir.empty(); }

    float deg2rad(float deg) { return (M_PI * deg) / 180.0; }
    float rad2deg(float rad) { return (180.0 * rad) / M_PI; }

    float getMinCartPos() { return leftTrackBound; }
    float getMaxCartPos() { return rightTrackBound; }
    float getMinCartVel() { return -3.0; }
    float getMaxCartVel() { return 3.0; }
    float getMinPoleAngle() { return leftAngleBound; }
    float getMaxPoleAngle() { return rightAngleBound; }
    float getMinPoleVelocity() { return minPoleVelocity; }
    float getMaxPoleVelocity() { return maxPoleVelocity; }
    float getMinPoleAccel() { return minPoleAccel; }
    float getMaxPoleAccel() { return maxPoleAccel; }
    float getPoleAngle() { return theta; }
    float getPoleAngleDegrees() { return rad2deg(theta); }
    float getPoleVelocity() { return theta_dot; }
    float getPoleAccel() { return theta_dd; }
    float getCartPosition() { return x; }
    float getCartVelocity() { return x_dot; }
    int getTimeAloft() { return timeAloft; }
    bool isInTimeout() { return timeoutCnt < timeoutDuration; }
    float getNetForce() { return netForce; }
    bool getErrorLeft() { return errorLeft; }
    bool getErrorRight() { return errorRight; }
    int getTrialNum() { return trialNum; }
    long getCycle() { return cycle; }
    bool getFallen() { return fallen; }
    float getForceLeft() { return forceLeft; }
    float getForceRight() { return forceRight; }
    bool getClosedLoop() { return closedLoop; }
    std::string getFailureMode();
    Microzone &leftMZ() { return forceLeftMZ; }
    Microzone &rightMZ() { return forceRightMZ; }

    static constexpr float gravity = 9.8f;
    static constexpr float cartMass = 1.0f;
    static constexpr float poleMass = 0.1f;
    static constexpr float length = 7.5f;  // half the pole's length
    static constexpr float maxForce = 1;
    static constexpr float minForce = -1;
    static constexpr float forceScale = .1f;
    static constexpr float forceDecay = .99f;
    static constexpr float tau = 0.001f;  // seconds per step
    static constexpr int timeoutDuration = 1500;
    static constexpr float leftAngleBound = -.1745f;  // 10 degrees
    static constexpr float rightAngleBound = .1745f;
    static constexpr float minPoleVelocity = -0.5f;
    static constexpr float maxPoleVelocity = 0.5f;
    static constexpr float minPoleAccel = -5.0f;
    static constexpr float maxPoleAccel = 5.0f;

    static float logScale(float value, float gain);
    static float inverseLogScale(float scaledVal, float gain);

protected:
    long maxTrialLength;
    int maxNumTrials;

    void init(const CartpoleOptions &opts);
    void reset();
    bool inFailure();
    float calcForce();
    void setMZErr();
    void computePhysics(float force);
    void advance(CBMSimCore *simCore, const ct_uint8_t *ncL, const ct_uint8_t *ncR);
    void saveState(CBMSimCore *simCore, const char *file);

    std::string saveStateDir;
    std::ofstream myfile;
    bool closedLoop, verbose;
    float trackLen, leftTrackBound, rightTrackBound;
    float polemassLength, totalMass;
    int timeoutCnt;
    int trialNum;
    int timeAloft;
    bool successful;
    int numMillionStepTrials;

    float x, x_dot, theta, old_theta, theta_dot, oldTheta_dot, theta_dd;

    Microzone forceLeftMZ, forceRightMZ;
    StateVariable<Cartpole> sv_highFreq, sv_poleVel, sv_poleAng, sv_cartVel, sv_cartPos;

    float netForce;
    float forceLeft, forceRight;
    bool errorLeft, errorRight;  // error delivered to the microzone this step
    bool fallen;
    long cycle;

    static const bool randomizeMFs = true;
    static constexpr float gaussWidth = 6.0f;
    static constexpr float highFreqMFProportion = .03f;
    static constexpr float poleAngMFProportion = .06f;
    static constexpr float poleVelMFProportion = .06f;
    static constexpr float cartPosMFProportion = .06f;
    static constexpr float cartVelMFProportion = .06f;
};

#endif
