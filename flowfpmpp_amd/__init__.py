"""MI355X-native stair-step influence-potential solver and FPM force densities."""
