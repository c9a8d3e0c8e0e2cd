// cartpole.cpp -- see CBMEnvInclude/environments/cartpole.hpp.
//
// The physics, error rules and trial bookkeeping follow
// Matthew_Sim_CBM/src/environments/cartpole.cpp:24-285 step for step, in
// binary32 with the reference's promotions: the pole's angle and velocity
// draws at reset and the angle wrap by 2*pi are binary64 rounded to binary32;
// cos/sin of the angle are the binary32 functions; the error draws compare a
// binary64 uniform with a binary32 probability.
#include <CBMEnvInclude/environments/cartpole.hpp>

#include <cfloat>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <stdexcept>

namespace {

const char *argValue(int argc, char **argv, const char *key)
{
    for (int i = 1; i + 1 < argc; i++)
        if (std::strcmp(argv[i], key) == 0) return argv[i + 1];
    return nullptr;
}

bool argFlag(int argc, char **argv, const char *key)
{
    for (int i = 1; i < argc; i++)
        if (std::strcmp(argv[i], key) == 0) return true;
    return false;
}

}  // namespace

Cartpole::Cartpole(CRandomSFMT0 *randGen, const CartpoleOptions &opts)
    : Environment(randGen), trialNum(-1), forceLeftMZ("ForceLeft", 0, forceScale, 1, forceDecay),
      forceRightMZ("ForceRight", 1, forceScale, 1, forceDecay), sv_highFreq("highFreqMFs", HIGH_FREQ, highFreqMFProportion),
      sv_poleVel("poleVelMFs", GAUSSIAN, poleVelMFProportion), sv_poleAng("poleAngMFs", GAUSSIAN, poleAngMFProportion),
      sv_cartVel("cartVelMFs", GAUSSIAN, cartVelMFProportion), sv_cartPos("cartPosMFs", GAUSSIAN, cartPosMFProportion)
{
    init(opts);
}

Cartpole::Cartpole(CRandomSFMT0 *randGen, int argc, char **argv)
    : Environment(randGen), trialNum(-1), forceLeftMZ("ForceLeft", 0, forceScale, 1, forceDecay),
      forceRightMZ("ForceRight", 1, forceScale, 1, forceDecay), sv_highFreq("highFreqMFs", HIGH_FREQ, highFreqMFProportion),
      sv_poleVel("poleVelMFs", GAUSSIAN, poleVelMFProportion), sv_poleAng("poleAngMFs", GAUSSIAN, poleAngMFProportion),
      sv_cartVel("cartVelMFs", GAUSSIAN, cartVelMFProportion), sv_cartPos("cartPosMFs", GAUSSIAN, cartPosMFProportion)
{
    // the defaults of Cartpole::getOptions (cartpole.cpp:12-22)
    CartpoleOptions o;
    o.logfile = "cartpole.log";
    o.simStateDir = "./";
    o.verbose = true;
    if (const char *v = argValue(argc, argv, "--logfile")) o.logfile = v;
    if (const char *v = argValue(argc, argv, "--maxNumTrials")) o.maxNumTrials = std::atoi(v);
    if (const char *v = argValue(argc, argv, "--maxTrialLen")) o.maxTrialLen = std::atol(v);
    if (const char *v = argValue(argc, argv, "--trackLen")) o.trackLen = (float)std::atof(v);
    if (const char *v = argValue(argc, argv, "--simStateDir")) o.simStateDir = v;
    o.closedLoop = argFlag(argc, argv, "--closedLoop");
    init(o);
}

void Cartpole::init(const CartpoleOptions &opts)
{
    saveStateDir = opts.simStateDir;
    maxTrialLength = opts.maxTrialLen;
    maxNumTrials = opts.maxNumTrials;
    closedLoop = opts.closedLoop;
    verbose = opts.verbose;
    numMillionStepTrials = 0;
    cycle = 0;
    timeAloft = 0;
    successful = false;
    x = x_dot = theta = old_theta = theta_dot = oldTheta_dot = theta_dd = 0;
    netForce = forceLeft = forceRight = 0;
    errorLeft = errorRight = false;
    fallen = false;
    timeoutCnt = 0;

    totalMass = cartMass + poleMass;
    polemassLength = poleMass * length;
    trackLen = opts.trackLen;
    if (trackLen <= 0) trackLen = std::numeric_limits<float>::max();
    leftTrackBound = -trackLen / 2.0;
    rightTrackBound = trackLen / 2.0;

    if (!opts.logfile.empty()) {
        myfile.open(opts.logfile.c_str());
        if (!myfile) throw std::runtime_error("Cartpole: cannot open log file " + opts.logfile);
        myfile << cycle << " TrackLen " << trackLen << " PoleLen " << 2 * length << " LeftAngleBound "
               << leftAngleBound << " RightAngleBound " << rightAngleBound << std::endl;
    }
    reset();

    stateVariables.push_back(&sv_highFreq);
    stateVariables.push_back(&sv_poleVel);
    stateVariables.push_back(&sv_poleAng);
    stateVariables.push_back(&sv_cartVel);
    stateVariables.push_back(&sv_cartPos);
    microzones.push_back(&forceLeftMZ);
    microzones.push_back(&forceRightMZ);
}

Cartpole::~Cartpole()
{
    if (myfile.is_open()) myfile.close();
}

void Cartpole::reset()
{
    if (myfile.is_open()) {
        myfile << cycle << " EndTrial " << trialNum << " TimeAloft " << timeAloft << " Failure: " << getFailureMode()
               << std::endl;
    }
    if (verbose)
        std::cout << "Trial " << trialNum << ": Time Aloft: " << timeAloft << " Failure: " << getFailureMode()
                  << std::endl;
    fallen = false;
    timeoutCnt = 0;
    x = 0.0;
    x_dot = 0.0;
    // a tiny random tilt and angular velocity (binary64, rounded)
    theta = (randGen->Random() - .5) * .00001;
    theta_dot = (randGen->Random() - .5) * .00001;
    old_theta = theta;
    oldTheta_dot = theta_dot;
    theta_dd = 0.0;
    trialNum++;
    successful = timeAloft >= maxTrialLength;
    timeAloft = 0;
}

void Cartpole::setupMossyFibers(CBMState *simState)
{
    setupMossyFibers((int)simState->sizes().numMF, (int)simState->sizes().numNC);
}

void Cartpole::setupMossyFibers(int nMF, int nNC)
{
    Environment::setupMossyFibers(nMF, nNC);
    Environment::setupStateVariables(randomizeMFs, myfile);
    sv_poleVel.initializeGaussian(minPoleVelocity, maxPoleVelocity, this, &Cartpole::getPoleVelocity);
    sv_poleAng.initializeGaussian(leftAngleBound, rightAngleBound, this, &Cartpole::getPoleAngle);
    sv_cartVel.initializeGaussian(getMinCartVel(), getMaxCartVel(), this, &Cartpole::getCartVelocity);
    sv_cartPos.initializeGaussian(getMinCartPos(), getMaxCartPos(), this, &Cartpole::getCartPosition);
}

float *Cartpole::getState()
{
    for (int i = 0; i < numMF; i++) mfFreq[i] = mfFreqRelaxed[i];
    // the always-on group at every step, the encoders only outside the time-out
    for (StateVariableBase *sv : stateVariables)
        if (sv->type == HIGH_FREQ || !isInTimeout()) sv->update();
    return mfFreq.data();
}

float Cartpole::logScale(float value, float gain)
{
    value *= gain;
    if (value >= 0) {
        value = std::max(1.0f, value);
        return std::log(value);
    }
    value = std::max(1.0f, -value);
    return -std::log(value);
}

float Cartpole::inverseLogScale(float scaledVal, float gain)
{
    if (scaledVal >= 0) return std::exp(scaledVal) / gain;
    return -std::exp(-scaledVal) / gain;
}

void Cartpole::step(CBMSimCore *simCore) { advance(simCore, nullptr, nullptr); }

void Cartpole::stepWithSpikes(CBMSimCore *simCore, const ct_uint8_t *apNCLeft, const ct_uint8_t *apNCRight)
{
    advance(simCore, apNCLeft, apNCRight);
}

void Cartpole::stepDetached(const ct_uint8_t *apNCLeft, const ct_uint8_t *apNCRight)
{
    advance(nullptr, apNCLeft, apNCRight);
}

void Cartpole::saveState(CBMSimCore *simCore, const char *file)
{
    if (saveStateDir.empty() || !simCore) return;
    std::string p = saveStateDir;
    if (p.back() != '/') p += '/';
    p += file;
    std::fstream f(p.c_str(), std::fstream::out | std::fstream::binary);
    if (!f) throw std::runtime_error("Cartpole: cannot write " + p);
    simCore->writeToState(f);
}

void Cartpole::advance(CBMSimCore *simCore, const ct_uint8_t *ncL, const ct_uint8_t *ncR)
{
    if (simCore) {
        if (!forceLeftMZ.initialized()) forceLeftMZ.initialize(simCore, numNC);
        if (!forceRightMZ.initialized()) forceRightMZ.initialize(simCore, numNC);
    } else {
        forceLeftMZ.initialize(NULL, numNC);
        forceRightMZ.initialize(NULL, numNC);
    }
    cycle++;
    timeoutCnt++;
    errorLeft = errorRight = false;
    forceLeftMZ.takeErrors();
    forceRightMZ.takeErrors();

    if (closedLoop) {
        if (ncL && ncR) {
            forceLeftMZ.update(ncL);
            forceRightMZ.update(ncR);
        } else if (simCore) {
            forceLeftMZ.update();
            forceRightMZ.update();
        } else {
            throw std::runtime_error("Cartpole: a closed-loop step needs NC spikes or a core");
        }
    }

    // State files around trials that reach the maximum length, and the end of
    // the run after a failure that follows five or more of them
    // (cartpole.cpp:153-176; the counter is a function static there).
    if (fallen && getFailureMode() == "MaxTrialLength") {
        saveState(simCore, numMillionStepTrials == 0 ? "milStepStart.st" : "milStepEnd.st");
        numMillionStepTrials++;
    } else if (fallen && getFailureMode() != "MaxTrialLength") {
        if (numMillionStepTrials >= 5) {
            saveState(simCore, "failureState.st");
            maxNumTrials = trialNum;
        }
        numMillionStepTrials = 0;
    }

    if (fallen) reset();

    if (timeoutCnt <= timeoutDuration) return;
    if (timeoutCnt == timeoutDuration + 1 && myfile.is_open())
        myfile << cycle << " StartingTrial " << trialNum << std::endl;

    computePhysics(calcForce());
    setMZErr();
    errorLeft = forceLeftMZ.takeErrors() > 0;
    errorRight = forceRightMZ.takeErrors() > 0;

    fallen = inFailure();
    timeAloft++;

    if (myfile.is_open())
        myfile << cycle << " Theta " << theta << " ThetaDot " << theta_dot << " CartPos " << x << " CartVel " << x_dot
               << " ForceLeft " << forceLeft << " ForceRight " << forceRight << " ErrorLeft " << errorLeft
               << " ErrorRight " << errorRight << " TimeAloft " << timeAloft << std::endl;
}

void Cartpole::computePhysics(float force)
{
    const float c = std::cos(theta);
    const float s = std::sin(theta);
    // angular and linear accelerations of the pole on the cart
    const float angNum = -1 * force * c - gravity * poleMass * s - gravity * cartMass * s +
                         length * poleMass * theta_dot * theta_dot * c * s;
    const float angDen = length * (poleMass + cartMass - poleMass * c * c);
    const float thetaacc = -angNum / angDen;
    const float linNum = -gravity * length * poleMass * c * s +
                         length * (-1 * force + length * poleMass * theta_dot * theta_dot * s);
    const float linDen = length * (cartMass + poleMass) - length * poleMass * c;
    const float xacc = -linNum / linDen;

    // Euler step of the four state variables
    x += tau * x_dot;
    x_dot += tau * xacc;
    theta += tau * theta_dot;
    theta_dot += tau * thetaacc;
    theta_dot = std::min(std::max(minPoleVelocity, theta_dot), maxPoleVelocity);
    theta_dd = std::min(std::max(minPoleAccel, thetaacc), maxPoleAccel);

    while (theta >= M_PI) theta -= 2.0 * M_PI;
    while (theta < -M_PI) theta += 2.0 * M_PI;
}

void Cartpole::setMZErr()
{
    const float maxErrProb = .01f;
    // the pole's angle: push against the fall
    float p = std::min(std::fabs(theta), maxErrProb);
    bool hit = randGen->Random() < p;
    if (theta >= 0.0 && theta_dot >= 0.0 && hit) forceLeftMZ.deliverError();
    if (theta < 0.0 && theta_dot < 0.0 && hit) forceRightMZ.deliverError();

    if (trackLen < std::numeric_limits<float>::max()) {
        // the cart's position: back towards the centre
        p = std::min(.005f * std::fabs(x) / rightTrackBound, maxErrProb);
        hit = randGen->Random() < p;
        if (x < 0 && hit) forceLeftMZ.deliverError();
        if (x > 0 && hit) forceRightMZ.deliverError();
        // the cart's velocity: both cases go to the left microzone in the
        // reference (cartpole.cpp:244-250), kept as it is
        p = std::min(.005f * std::fabs(x_dot), maxErrProb);
        hit = randGen->Random() < p;
        if (x_dot > 0 && x < 0 && hit) forceLeftMZ.deliverError();
        if (x_dot < 0 && x > 0 && hit) forceLeftMZ.deliverError();
    }
}

bool Cartpole::inFailure()
{
    return theta <= leftAngleBound || theta >= rightAngleBound || x <= leftTrackBound || x >= rightTrackBound ||
           timeAloft >= maxTrialLength;
}

std::string Cartpole::getFailureMode()
{
    if (theta <= leftAngleBound) return "PoleFallRight";
    if (theta >= rightAngleBound) return "PoleFallLeft";
    if (x <= leftTrackBound) return "CartFallLeft";
    if (x >= rightTrackBound) return "CartFallRight";
    if (timeAloft >= maxTrialLength) return "MaxTrialLength";
    return "NoFailure";
}

float Cartpole::calcForce()
{
    forceRight = forceRightMZ.getForce();
    forceLeft = forceLeftMZ.getForce();
    netForce = forceRight - forceLeft;
    return netForce;
}

bool Cartpole::terminated() { return trialNum >= maxNumTrials; }
