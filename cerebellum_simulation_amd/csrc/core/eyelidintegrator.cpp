// eyelidintegrator.cpp -- see CBMToolsInclude/eyelidintegrator.h.  Binary32
// arithmetic in the reference's operand order (eyelidintegrator.cpp:52-68).
#include <CBMToolsInclude/eyelidintegrator.h>

#include <cmath>

EyelidIntegrator::EyelidIntegrator(unsigned int numInputCells, unsigned int msPerTimeStep, float gDecayTau,
                                   float gIncrease, float gShift, float gLeakRaw, float maxAmplitude)
    : g(numInputCells, 0.0f), inc(gIncrease), shift(gShift), amp(maxAmplitude), pos(0.0f)
{
    const int ms = (int)msPerTimeStep;
    decay = std::exp(-ms / gDecayTau);  // float argument: the float overload
    leak = gLeakRaw / (6 - ms);
}

float EyelidIntegrator::calcStep(const ct_uint8_t *apIn)
{
    float sum = 0;
    for (size_t i = 0; i < g.size(); i++) {
        g[i] = g[i] * decay;
        g[i] = g[i] + apIn[i] * inc;
        sum += g[i];
    }
    sum = sum + shift;
    sum = (sum > 0) * sum;
    pos = pos + (leak * (0 - pos)) + (sum * (amp - pos));
    return pos;
}
