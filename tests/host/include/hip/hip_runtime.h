// Empty stand-in for HIP's runtime header so that the device helper headers
// (plain integer code) compile for the host emulation tests in tests/host/.
#pragma once
