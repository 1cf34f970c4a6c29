// Test infrastructure only. The colour-matrix stage of the reference's
// SaveRender (src/pathtracer.cpp:3501-3503) evaluated with the glm library the
// reference vendors (includes/glm): Bradford adaptation E -> D65, then XYZ ->
// linear sRGB, each as `XYZ * glm::mat3(...)` (src/pathtracer.cpp:970-986), then
// glm::max with zero. The matrices are the published ones (Lindbloom,
// brucelindbloom.com); glm::mat3's constructor takes them column by column and
// `vec3 * mat3` dots the vector with each column, so each written row of nine
// numbers acts as a row of the matrix.
//
// Reads float32 triples from stdin and writes the resulting float32 triples to
// stdout. Built with -ffp-contract=off, so each product and sum is rounded as
// written. Built by oracle/Makefile into oracle/_ref/ from the headers where
// they lie in the reference tree; never part of the product.
#include <cstdio>

#include <glm/glm.hpp>

int main() {
    const glm::mat3 bradford(0.9531874, -0.0265906, 0.0238731,
                             -0.0382467, 1.0288406, 0.0094060,
                             0.0026068, -0.0030332, 1.0892565);
    const glm::mat3 to_srgb(3.2404542, -1.5371385, -0.4985314,
                            -0.9692660, 1.8760108, 0.0415560,
                            0.0556434, -0.2040259, 1.0572252);
    float v[3];
    while (std::fread(v, sizeof(float), 3, stdin) == 3) {
        const glm::vec3 xyz(v[0], v[1], v[2]);
        const glm::vec3 rgb = glm::max((xyz * bradford) * to_srgb, glm::vec3(0.0f));
        const float out[3] = {rgb.x, rgb.y, rgb.z};
        std::fwrite(out, sizeof(float), 3, stdout);
    }
    return 0;
}
