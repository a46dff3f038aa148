"""phaser_amd -- MI355X-native correlation core for PHASER-style registration."""
