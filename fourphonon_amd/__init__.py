"""MI355X-native four-phonon scattering hot path (FourPhonon drop-in)."""
