"""MI355X-native (gfx950) Fock-build hot path of QUICK: Schwarz-screened s/p/d
ERIs contracted with the density into the Coulomb/exchange operator."""
