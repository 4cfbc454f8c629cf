// What the fake bake of fake_device.cpp writes, for the server driver to check results against.
#pragma once
#include <cstdint>

uint64_t fake_soup_key(const float *xyz, uint64_t ntri);
float fake_expected_distance(uint64_t key, uint64_t i);
bool fake_expected_brick(uint64_t key, uint64_t b);
uint32_t fake_expected_word(uint64_t key, uint64_t i);
uint64_t fake_mesh_bytes(uint64_t ntri);
