// Umbrella header of the MI355X IANE convolution / deconvolution drop-in.
#pragma once

#include <iane/core.hpp>
#include <iane/convolution.hpp>
#include <iane/prepared.hpp>
#include <iane/deconvolution.hpp>
#include <iane/run_on.hpp>
#include <iane/expression.hpp>
#include <iane/texture.hpp>
#include <iane/sequence.hpp>
