// bvh.cpp — host-built bounding volume hierarchy for the extended mode.
